// SPH step kernels for gfx950: kick/drift/key, cell ranges, sorted gather,
// density (neighbour-list build + dense accumulation), fused force + kick,
// adaptive refinement (criteria, merge pairing, merge, compaction, split) and
// the diagnostics. All kernels are segmented over a scene table, so one launch
// steps every scene of a batch; a single simulation is a batch of one.
//
// Reference: cuda/src/algorithm/{Algorithm,BoundaryAlgorithm,AdaptiveAlgorithm}.cu,
// cuda/src/algorithm/neighbors/NeighborKernels.cu,
// cuda/src/simulation/neighbors/NeighborGrid.cuh,
// cuda/src/simulation/refinement/criteria/*.cuh.
#include <cfloat>

#include "kernels.hpp"

namespace asph
{

namespace
{

constexpr uint32_t kGridCap = 16384;

__host__ __forceinline__ uint32_t gridFor(uint32_t n, uint32_t block)
{
    const uint32_t g = (n + block - 1) / block;
    return g == 0 ? 1U : (g > kGridCap ? kGridCap : g);
}

// A grid of one thread per slot, uncapped, for kernels without a grid-stride
// loop.
__host__ __forceinline__ uint32_t gridCover(uint32_t n, uint32_t block)
{
    const uint32_t g = static_cast<uint32_t>((static_cast<uint64_t>(n) + block - 1U) / block);
    return g == 0 ? 1U : g;
}

__device__ uint32_t blockScan256(uint32_t v, uint32_t* lds, uint32_t* total)
{
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = tid >> 6;
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
    uint32_t base = 0;
    uint32_t sum = 0;
    for (int w = 0; w < 4; ++w)
    {
        const uint32_t t = lds[w];
        if (w < wave)
        {
            base += t;
        }
        sum += t;
    }
    __syncthreads();
    *total = sum;
    return base + x - v;
}

__device__ __forceinline__ bool mergeParticipant(const SceneDev& sc, const SceneState& st)
{
    return (sc.flags & kSceneActive) != 0U && st.refineNow != 0U && st.countBefore > 1U;
}

// ------------------------------------------------------------------ prep --
// Per step: dense offsets of the scenes, simulated time, sloshing gravity,
// refinement-zone centre and the refinement cadence
// (AdaptiveSphSimulation.cu:75-79, frame counter incremented at :84).
__global__ __launch_bounds__(256) void kPrep(Ctx c)
{
    __shared__ uint32_t lds[4];
    __shared__ uint32_t anyQuery;
    if (threadIdx.x == 0)
    {
        anyQuery = 0U;
    }
    __syncthreads();
    uint32_t anyOrdered = 0;
    uint32_t carry = 0;
    for (uint32_t s0 = 0; s0 < c.sceneSlots; s0 += 256)
    {
        const uint32_t s = s0 + threadIdx.x;
        bool active = false;
        uint32_t cnt = 0;
        if (s < c.sceneSlots)
        {
            active = (c.scenes[s].flags & kSceneActive) != 0U;
            cnt = active ? c.states[s].count : 0U;
        }
        uint32_t tot = 0;
        const uint32_t ex = blockScan256(cnt, lds, &tot);
        if (s < c.sceneSlots)
        {
            c.denseStarts[s] = carry + ex;
            SceneState& st = c.states[s];
            st.denseOff = carry + ex;
            if (active)
            {
                const SceneDev& sc = c.scenes[s];
                const double t = st.time + static_cast<double>(sc.dt);
                float gx = sc.g[0];
                if (sc.sloshA != 0.0F && sc.sloshT != 0.0F)
                {
                    double sn = 0.0;
                    double cs = 0.0;
                    detSinCos(6.283185307179586 * t / static_cast<double>(sc.sloshT), &sn, &cs);
                    gx = static_cast<float>(static_cast<double>(sc.g[0]) - static_cast<double>(sc.sloshA) * sn);
                }
                st.gEff[0] = gx;
                st.gEff[1] = sc.g[1];
                st.gEff[2] = sc.g[2];
                st.zoneNow[0] = sc.zoneC[0];
                st.zoneNow[1] = sc.zoneC[1];
                st.zoneNow[2] = sc.zoneC[2];
                if (sc.zonePathT != 0.0F)
                {
                    double sn = 0.0;
                    double cs = 0.0;
                    detSinCos(6.283185307179586 * t / static_cast<double>(sc.zonePathT), &sn, &cs);
                    st.zoneNow[0] =
                        static_cast<float>(static_cast<double>(sc.zoneC[0]) + static_cast<double>(sc.zonePathR) * cs);
                    st.zoneNow[2] =
                        static_cast<float>(static_cast<double>(sc.zoneC[2]) + static_cast<double>(sc.zonePathR) * sn);
                }
                uint32_t refine = 0;
                if ((sc.flags & kSceneAdaptive) != 0U)
                {
                    const uint32_t f = st.frame;
                    const uint32_t ic = sc.initialCooldown;
                    refine = (f == ic || (f > ic && (f % sc.cooldown) == 0U)) ? 1U : 0U;
                    st.frame = f + 1U;
                }
                st.refineNow = refine;
                st.time = t;
                if (c.queryTileBits == kQueryAuto && st.qorder == 0U && st.qhits >= kQueryMinHits)
                {
                    // slabbed queries once the sampled waves kept more than
                    // 8.5 survivors per neighbour of a particle (a freshly
                    // generated lattice: about 6.5; a mixed dam break: 14);
                    // the counts accumulate over steps until they cover
                    // enough waves
                    const uint64_t surv = st.qsurv;
                    const uint64_t hits = st.qhits;
                    st.qorder = surv * 128U > hits * 17U ? 1U : 0U;
                    st.qsurv = 0U;
                    st.qhits = 0U;
                }
                if (st.qorder != 0U || c.queryTileBits != kQueryAuto)
                {
                    anyOrdered = 1U;
                }
                st.positives = 0;
                st.mergePairs = 0;
                if (refine != 0U)
                {
                    st.refined = 1U;
                    st.countBefore = cnt;
                    st.mergeK = 0;
                    st.removed = 0;
                    st.countAfterMerge = cnt;
                    st.splitK = 0;
                    st.splitDone = 0;
                    st.countAfter = cnt;
                }
                else
                {
                    st.refined = 0U;
                }
            }
        }
        carry += tot;
    }
    // merge layout: the criterion keys of the merging scenes back to back, in
    // scene order, from each scene's segStart (findTopParticlesToMerge sorts
    // only these)
    uint32_t mcarry = 0;
    for (uint32_t s0 = 0; s0 < c.sceneSlots; s0 += 256)
    {
        const uint32_t s = s0 + threadIdx.x;
        uint32_t mc = 0;
        if (s < c.sceneSlots && mergeParticipant(c.scenes[s], c.states[s]))
        {
            mc = c.states[s].countBefore;
        }
        uint32_t tot = 0;
        const uint32_t ex = blockScan256(mc, lds, &tot);
        if (mc > 0U)
        {
            c.states[s].segStart = mcarry + ex;
        }
        mcarry += tot;
    }
    if (anyOrdered != 0U)
    {
        atomicOr(&anyQuery, 1U);
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        *c.queryAny = anyQuery;
        // heavy-wave threshold from the last step's sampled costs
        const uint32_t sampled = c.wstat[1];
        c.wstat[2] = sampled >= 8U ? static_cast<uint32_t>((static_cast<uint64_t>(c.wstat[0]) * kHeavyFactor2) /
                                                            (2U * sampled))
                                   : 0U;
        c.wstat[0] = 0U;
        c.wstat[1] = 0U;
        *c.heavyCount = 0U;
        *c.nTotal = carry;
        *c.overflowCount = 0U;
        *c.wallQueueCount = 0U;
        if (c.critCount != nullptr)
        {
            *c.critCount = mcarry;
        }
    }
}

// ------------------------------------------------- kick + drift + key --
// halfKickVelocities (Algorithm.cu:270-284), updatePositions (:286-298),
// assignParticlesToCells (NeighborKernels.cu:5-15) and the cell histogram.
__global__ __launch_bounds__(1024) void kKickDriftKey(Ctx c)
{
    const uint32_t n = *c.nTotal;
    for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < n; d += gridDim.x * blockDim.x)
    {
        const uint32_t s = segmentOf(c.denseStarts, c.sceneSlots, d);
        const SceneDev& sc = c.scenes[s];
        const uint32_t i = sc.pOff + (d - c.denseStarts[s]);
        F4 v = c.vel[i];
        const F4 a = c.acc[i];
        const float hdt = sc.halfDt;
        v.x = v.x + a.x * hdt;
        v.y = v.y + a.y * hdt;
        v.z = v.z + a.z * hdt;
        v.w = v.w + a.w * hdt;
        const float len = sqrtf(dot4(v.x, v.y, v.z, v.w));
        if (len > sc.vmax)
        {
            const float k = sc.vmax / len;
            v.x = v.x * k;
            v.y = v.y * k;
            v.z = v.z * k;
            v.w = v.w * k;
        }
        F4 p = c.pos[i];
        const float dt = sc.dt;
        p.x = p.x + v.x * dt;
        p.y = p.y + v.y * dt;
        p.z = p.z + v.z * dt;
        p.w = p.w + v.w * dt;
        p.x = gmin(gmax(p.x, sc.dmin[0]), sc.dmax[0]);
        p.y = gmin(gmax(p.y, sc.dmin[1]), sc.dmax[1]);
        p.z = gmin(gmax(p.z, sc.dmin[2]), sc.dmax[2]);
        c.vel[i] = v;
        c.pos[i] = p;
        const uint32_t key = sc.cellBase + cellKey(sc, p.x, p.y, p.z);
        c.keysA[d] = key;
        c.valsA[d] = i;
    }
}

// ------------------------------------------------------- scene ranges --
// Sorted range of each scene and the slot-0 omission of
// calculateCellStartAndEndIndices (NeighborKernels.cu:27-36): when the
// lowest-key cell of a scene holds exactly one particle, that cell and the
// next occupied cell are invisible to every neighbour walk.
// First sorted slot whose key is at least `key` (keys ascending over [0, n)).
__device__ __forceinline__ uint32_t lowerBound(const uint32_t* keys, uint32_t n, uint32_t key)
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

// Per scene in cell-base order: its first sorted slot, and the end of the
// slots that the reference's cell start/end construction hides from every
// walk (NeighborKernels.cu:27-36: with the slot-0 omission, the lowest-key
// cell is lost if it holds one particle, together with the next cell).
__device__ void sceneRanges(const Ctx& c, uint32_t n, uint32_t nOrder)
{
    for (uint32_t k = threadIdx.x; k < nOrder; k += blockDim.x)
    {
        const uint32_t s = c.cellOrder[k];
        const SceneDev& sc = c.scenes[s];
        SceneState& st = c.states[s];
        const uint32_t ss = lowerBound(c.sortedKeys, n, sc.cellBase);
        c.sortedStarts[k] = ss;
        st.sortedStart = ss;
        const uint32_t cnt = st.count;
        uint32_t lost = ss;
        if ((sc.flags & kSceneFixCellStart) == 0U && cnt > 0U)
        {
            if (cnt == 1U)
            {
                lost = ss + 1U;
            }
            else
            {
                const uint32_t k0 = c.sortedKeys[ss];
                const uint32_t k1 = c.sortedKeys[ss + 1U];
                if (k0 != k1)
                {
                    lost = lowerBound(c.sortedKeys, n, k1 + 1U);
                }
            }
        }
        st.lostEnd = lost;
    }
}

// Sorted shadows and group boxes; also the cell offsets of the sorted order
// (CSR: cellOff[c] = first slot whose cell is at least c, the exclusive scan
// of the cell histogram) and, in block 0, the per-scene ranges. The kernels
// after this one read both.
__global__ __launch_bounds__(1024) void kGather(Ctx c)
{
    const uint32_t n = *c.nTotal;
    const uint32_t nOrder = *c.nOrderPtr;
    const uint32_t nPad = (n + 63U) & ~63U;
    if (blockIdx.x == 0U)
    {
        sceneRanges(c, n, nOrder);
    }
    for (uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x; cell < c.cExtent; cell += gridDim.x * blockDim.x)
    {
        c.cellOff[cell] = lowerBound(c.sortedKeys, n, cell);
    }
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nPad; t += gridDim.x * blockDim.x)
    {
        float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX;
        float hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
        float mr = 0.0F;
        if (t < n)
        {
            const uint32_t k = segmentOf(c.cellBaseOrder, nOrder, c.sortedKeys[t]);
            const uint32_t i = c.sortedVals[t];
            const F4 p = c.pos[i];
            const F4 v = c.vel[i];
            const float hh = c.h[i];
            const float mm = c.m[i];
            const float r = 2.0F * hh;
            const float r2 = r * r;
            c.s0[t] = F4 {p.x, p.y, p.z, r2};
            c.s1[t] = F4 {v.x, v.y, v.z, mm};
            // the per-particle factors of wendlandKernel / wendlandDerivativeKernel
            // (WendlandKernel.cuh:39-52, 69-84), same expressions and rounding
            const float h3 = (hh * hh) * hh;
            const float h4 = h3 * hh;
            c.shm[t] = F4 {hh, mm, kWCoef / h3, kWDCoef / h4};
            const uint32_t s = c.cellOrder[k];
            c.sscene[t] = s;
            if ((c.scenes[s].flags & kSceneAdaptive) != 0U)
            {
                c.slotOf[i] = t;  // the merge search reads the particle's neighbour list
            }
            lx = hx = p.x;
            ly = hy = p.y;
            lz = hz = p.z;
            mr = r2;
        }
        for (int off = 1; off < static_cast<int>(kGroup); off <<= 1)
        {
            lx = fminf(lx, __shfl_xor(lx, off, 64));
            ly = fminf(ly, __shfl_xor(ly, off, 64));
            lz = fminf(lz, __shfl_xor(lz, off, 64));
            hx = fmaxf(hx, __shfl_xor(hx, off, 64));
            hy = fmaxf(hy, __shfl_xor(hy, off, 64));
            hz = fmaxf(hz, __shfl_xor(hz, off, 64));
            mr = fmaxf(mr, __shfl_xor(mr, off, 64));
        }
        if ((t & (kGroup - 1U)) == 0U && t < n)
        {
            c.gbox0[t / kGroup] = F4 {lx, ly, lz, mr};
            c.gbox1[t / kGroup] = F4 {hx, hy, hz, 0.0F};
        }
        if ((t & 63U) == 0U && t < n)
        {
            // the list builder's wave order: the waves of 64 query positions
            // that cost most in the last step are listed to run first
            const uint32_t w = t / 64U;
            const uint32_t cost = c.wcost[w];
            const uint32_t thr = c.wstat[2];
            uint32_t listed = 0U;
            if (cost == kCostPerLane || (thr != 0U && cost > thr))
            {
                const uint32_t k = atomicAdd(c.heavyCount, 1U);
                if (k < kHeavyCap)
                {
                    c.heavyList[k] = w;
                    listed = 1U;
                }
            }
            c.wheavy[w] = listed;
        }
    }
}

// True unless every member of the group is provably farther from q than its
// own support 2h (conservative by a relative margin far above FP32 rounding).
__device__ __forceinline__ bool groupMayHit(const F4& b0, const F4& b1, const F4& q)
{
    const float ex = fmaxf(fmaxf(b0.x - q.x, q.x - b1.x), 0.0F);
    const float ey = fmaxf(fmaxf(b0.y - q.y, q.y - b1.y), 0.0F);
    const float ez = fmaxf(fmaxf(b0.z - q.z, q.z - b1.z), 0.0F);
    const float d2 = (ex * ex + ey * ey) + ez * ez;
    return d2 <= b0.w * 1.0001F;
}

// False only if every point of the box [b0, b1] is provably farther from q
// than sqrt(r2): the box distance is a lower bound of every member's
// distance, and the margin covers its FP32 rounding. An empty (inverted) box
// is never reached.
__device__ __forceinline__ bool boxMayReach(const F4& b0, const F4& b1, const F4& q, float r2)
{
    const float ex = fmaxf(fmaxf(b0.x - q.x, q.x - b1.x), 0.0F);
    const float ey = fmaxf(fmaxf(b0.y - q.y, q.y - b1.y), 0.0F);
    const float ez = fmaxf(fmaxf(b0.z - q.z, q.z - b1.z), 0.0F);
    const float d2 = (ex * ex + ey * ey) + ez * ez;
    return d2 <= r2 * 1.0001F;
}

// Calls f(j, o) for every slot j in [lo, hi) (ascending) with o = s0[j] and
// d2(q, o) <= o.w, visiting only groups that may hold such slots.
template <class F>
__device__ __forceinline__ void scanRange(const Ctx& c, uint32_t lo, uint32_t hi, const F4& q, F&& f)
{
    if (lo >= hi)
    {
        return;
    }
    for (uint32_t g = lo / kGroup; g <= (hi - 1U) / kGroup; ++g)
    {
        if (!groupMayHit(c.gbox0[g], c.gbox1[g], q))
        {
            continue;
        }
        // the group's members in batches of 4 loads (a small register footprint)
        for (uint32_t base = g * kGroup; base < (g + 1U) * kGroup; base += 4U)
        {
            F4 o[4];
#pragma unroll
            for (uint32_t u = 0; u < 4U; ++u)
            {
                o[u] = c.s0[min(base + u, hi - 1U)];
            }
#pragma unroll
            for (uint32_t u = 0; u < 4U; ++u)
            {
                const uint32_t j = base + u;
                if (j < lo || j >= hi)
                {
                    continue;
                }
                const float dx = o[u].x - q.x;
                const float dy = o[u].y - q.y;
                const float dz = o[u].z - q.z;
                const float d2 = dot3w(dx, dy, dz, dx, dy, dz);
                if (d2 <= o[u].w)
                {
                    f(j, o[u], d2);
                }
            }
        }
    }
}

// Fluid walk of forEachFluidNeighbor (NeighborGrid.cuh:90-126) restricted to
// accepted neighbours (d^2 <= (2 h_j)^2), in the reference's visiting order.
template <class F>
__device__ __forceinline__ void walkFluid(const Ctx& c, const SceneDev& sc, uint32_t lost, const CellBox& b,
                                          const F4& q, F&& f)
{
    const int gxgy = sc.gx * sc.gy;
    for (int x = b.x0; x <= b.x1; ++x)
    {
        for (int y = b.y0; y <= b.y1; ++y)
        {
            for (int z = b.z0; z <= b.z1; ++z)
            {
                const uint32_t cell = sc.cellBase + static_cast<uint32_t>(x + y * sc.gx + z * gxgy);
                scanRange(c, max(c.cellOff[cell], lost), c.cellOff[cell + 1U], q, f);
            }
        }
    }
}

// False if no boundary particle can lie within R of sorted slot t (the
// per-cell bound of the squared distance to the walls exceeds R^2).
__device__ __forceinline__ bool wallsMayReach(const Ctx& c, uint32_t t, float R)
{
    return R * R >= c.bnear2[c.sortedKeys[t]];
}

__device__ __forceinline__ bool touchesEdge(const SceneDev& sc, const CellBox& b)
{
    if ((sc.flags & kSceneBoundaryInside) != 0U)
    {
        return true;
    }
    return b.x0 == 0 || b.y0 == 0 || b.z0 == 0 || b.x1 == sc.gx - 1 || b.y1 == sc.gy - 1 || b.z1 == sc.gz - 1;
}

// ------------------------------------------------------- wave helpers --
__device__ __forceinline__ int waveMinI(int v)
{
    for (int off = 32; off > 0; off >>= 1)
    {
        v = min(v, __shfl_xor(v, off, 64));
    }
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ int waveMaxI(int v)
{
    for (int off = 32; off > 0; off >>= 1)
    {
        v = max(v, __shfl_xor(v, off, 64));
    }
    return __builtin_amdgcn_readfirstlane(v);
}


// Orders a wave's LDS writes before its lanes' later reads of them (the
// lanes of a wave execute together; this keeps the compiler from moving
// accesses across).
__device__ __forceinline__ void waveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bool queryOrdered(const Ctx& c, uint32_t s)
{
    return c.qperm != nullptr && (c.queryTileBits != kQueryAuto || c.states[s].qorder != 0U);
}

// The sorted slot whose neighbour list sits at query position pq (the query
// order permutes slots within cells, so pq and the slot share a scene).
__device__ __forceinline__ uint32_t querySlot(const Ctx& c, uint32_t pq)
{
    if (c.qperm == nullptr || (c.queryTileBits == kQueryAuto && c.states[c.sscene[pq]].qorder == 0U))
    {
        return pq;
    }
    return c.qperm[pq];
}

// Query order (LaunchCfg::queryTile): within each cell of more than 64
// slots, the slots grouped by their sub-cell, one wave per cell: modes 1 and
// 2 take 2x2x2 or 4x4x4 sub-cells in Morton order, mode 3 4x4x4 sub-cells in
// x-major order (x, then y, then z), so that any 64 consecutive queries form
// a slab of the cell across x and, as the next cell in key order is the next
// along x, a wave that runs into the next cell continues the same slab (the
// order a lattice generated with x slowest has in index order).
// Any grouping gives the same lists and sums (each lane's list depends on its
// own query only); grouping keeps a wave's queries close together, so its
// candidate filter keeps few survivors even in refined cells, whose particles
// are in index order (daughters appended per split), and in mixed fluid,
// where a cell's particles in index order lie all over it.

// The query order of one cell of more than 64 slots [lo, hi): a counting
// sort of its slots by sub-cell (64 bins in LDS), ties in any order.
__device__ __forceinline__ void queryOrderCell(const Ctx& c, uint32_t lo, uint32_t hi, uint32_t lane, uint32_t* bin)
{
    bin[lane] = 0U;
    const SceneDev& sc = c.scenes[c.sscene[lo]];
    const uint32_t mode = c.queryTileBits == kQueryAuto ? 3U : c.queryTileBits;  // 1, 2 or 3
    const uint32_t side = mode == 1U ? 2U : 4U;
    // the position within its cell in units of sub-cells (only the grouping
    // depends on it, so reciprocals serve)
    const float sx = static_cast<float>(side) / sc.cs[0];
    const float sy = static_cast<float>(side) / sc.cs[1];
    const float sz = static_cast<float>(side) / sc.cs[2];
    auto subOf = [&](const F4& q) {
        const float u[3] = {(q.x - sc.dmin[0]) * sx, (q.y - sc.dmin[1]) * sy, (q.z - sc.dmin[2]) * sz};
        uint32_t k[3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            const float f = u[a] - static_cast<float>(side) * floorf(u[a] / static_cast<float>(side));
            k[a] = min(static_cast<uint32_t>(fmaxf(f, 0.0F)), side - 1U);
        }
        if (mode == 3U)
        {
            return (k[0] * 4U + k[1]) * 4U + k[2];
        }
        uint32_t m = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
        {
            // Morton: bit b of axis a at 3b + a (below 64 bins)
            m |= ((k[a] & 1U) << a) | (((k[a] >> 1U) & 1U) << (a + 3));
        }
        return m;
    };
    // the first kKeep rounds of 64 slots: positions loaded together, their
    // sub-cells kept for the placement pass
    constexpr uint32_t kKeep = 4;
    uint32_t kept[kKeep];
    {
        F4 q[kKeep];
#pragma unroll
        for (uint32_t r = 0; r < kKeep; ++r)
        {
            const uint32_t t = lo + r * 64U + lane;
            q[r] = c.s0[t < hi ? t : lo];
        }
#pragma unroll
        for (uint32_t r = 0; r < kKeep; ++r)
        {
            kept[r] = subOf(q[r]);
        }
    }
    waveSync();
#pragma unroll
    for (uint32_t r = 0; r < kKeep; ++r)
    {
        if (lo + r * 64U + lane < hi)
        {
            atomicAdd(&bin[kept[r]], 1U);
        }
    }
    for (uint32_t t0 = lo + kKeep * 64U; t0 < hi; t0 += 64U)
    {
        const uint32_t t = t0 + lane;
        if (t < hi)
        {
            atomicAdd(&bin[subOf(c.s0[t])], 1U);
        }
    }
    waveSync();
    // exclusive scan of the 64 bins across the wave
    const uint32_t count = bin[lane];
    uint32_t incl = count;
    for (uint32_t off = 1; off < 64U; off <<= 1U)
    {
        const uint32_t y = __shfl_up(incl, off, 64);
        incl += lane >= off ? y : 0U;
    }
    bin[lane] = lo + incl - count;
    waveSync();
    for (uint32_t t0 = lo, r = 0; t0 < hi; t0 += 64U, ++r)
    {
        const uint32_t t = t0 + lane;
        if (t < hi)
        {
            uint32_t sub = 0U;
#pragma unroll
            for (uint32_t rr = 0; rr < kKeep; ++rr)
            {
                sub = r == rr ? kept[rr] : sub;
            }
            if (r >= kKeep)
            {
                sub = subOf(c.s0[t]);
            }
            const uint32_t p = atomicAdd(&bin[sub], 1U);
            c.qperm[p] = t;
            c.qpos[t] = p;
        }
    }
    waveSync();
}

__global__ __launch_bounds__(256) void kQueryOrder(Ctx c)
{
    __shared__ uint32_t bins[4][64];
    const uint32_t lane = threadIdx.x & 63U;
    const uint32_t w = threadIdx.x >> 6U;
    const uint32_t waves = gridDim.x * 4U;
    if (*c.queryAny == 0U)
    {
        return;  // every scene keeps the sorted order this step
    }
    // cellOff holds entries 0 .. cExtent - 1 (the last one an end entry):
    // the ranges of cells 0 .. cExtent - 2 cover every sorted slot once
    for (uint32_t cell = blockIdx.x * 4U + w; cell + 1U < c.cExtent; cell += waves)
    {
        const uint32_t lo = c.cellOff[cell];
        const uint32_t hi = c.cellOff[cell + 1U];
        if (hi <= lo || !queryOrdered(c, c.sscene[lo]))
        {
            continue;  // empty, or a scene that keeps the sorted order
        }
        if (hi - lo <= 64U)
        {
            // one wave's worth: the sorted order
            if (lane < hi - lo)
            {
                c.qperm[lo + lane] = lo + lane;
                c.qpos[lo + lane] = lo + lane;
            }
            continue;
        }
        queryOrderCell(c, lo, hi, lane, bins[w]);
    }
}

// Read-only view in the constant address space: loads with wave-uniform
// addresses then go through the scalar cache (the array is not written by the
// kernels that read it this way).
typedef const __attribute__((address_space(4))) uint32_t ConstU32;

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup i runs on
// XCD i mod 8), and each XCD has its own L2. The neighbour-list consumers
// renumber their workgroups so that each XCD takes runs of kForceRunWaves
// consecutive waves of the sorted particles, the runs dealt to the XCDs in
// turn: neighbouring particles (which gather the same neighbour data) then
// share an L2, while a costly region (refined cells) still spreads over every
// XCD. The runs are taken of the workgroups that have particles (the grid is
// sized for the pool's capacity, which can be many times n). Results do not
// depend on the numbering. Force stage by run length in waves (ms; 500k
// adaptive tank / 256-scene batch / 1M tolerance mode): dispatch order
// 0.58 / 8.70 / 0.248, 32: 0.54 / 8.28 / 0.246, 64: 0.54 / 8.33 / 0.245,
// 128: 0.55 / 8.42 / 0.242, 256: 0.65 / 8.67 / 0.243, one contiguous eighth
// per XCD 0.65 / 9.98 / 0.248. (kNeighbors deals its waves in longer runs:
// kNeighborRunWaves.)
#ifndef ASPH_FORCE_RUN_WAVES
#define ASPH_FORCE_RUN_WAVES 64
#endif
constexpr uint32_t kForceRunWaves = ASPH_FORCE_RUN_WAVES;
__device__ __forceinline__ uint32_t xcdBlock(uint32_t n)
{
    constexpr uint32_t kXcds = 8;
    const uint32_t needed = (n + blockDim.x - 1U) / blockDim.x;
    const uint32_t used = needed < gridDim.x ? needed : gridDim.x;
    const uint32_t b = blockIdx.x;
    const uint32_t G = kForceRunWaves / (blockDim.x / 64U) > 0U ? kForceRunWaves / (blockDim.x / 64U) : 1U;
    if (b >= (used / (kXcds * G)) * (kXcds * G))
    {
        return b;  // the tail past the last whole round of runs keeps its order
    }
    const uint32_t k = b / kXcds;
    return (k / G) * (kXcds * G) + (b % kXcds) * G + k % G;
}

// Neighbour-list entries whose data are loaded together before the pair math
// (memory-level parallelism against register footprint).
#ifndef ASPH_FORCE_BATCH
#define ASPH_FORCE_BATCH 4
#endif
#ifndef ASPH_FORCE_TOL_WAVES
#define ASPH_FORCE_TOL_WAVES 4
#endif
#ifndef ASPH_FORCE_TOL_BATCH
#define ASPH_FORCE_TOL_BATCH 8
#endif
#ifndef ASPH_DENSITY_UNIFORM_BATCH
#define ASPH_DENSITY_UNIFORM_BATCH 8
#endif
constexpr uint32_t kListDensityUniformBatch = ASPH_DENSITY_UNIFORM_BATCH;  // a multiple of 4
#ifndef ASPH_LIST_DENSITY_BATCH
#define ASPH_LIST_DENSITY_BATCH 4
#endif
constexpr uint32_t kListDensityBatch = ASPH_LIST_DENSITY_BATCH;  // a multiple of 4
constexpr uint32_t kForceBatch = ASPH_FORCE_BATCH;      // a multiple of 4
#ifndef ASPH_FORCE_UNIFORM_BATCH
#define ASPH_FORCE_UNIFORM_BATCH 8
#endif
constexpr uint32_t kForceUniformBatch = ASPH_FORCE_UNIFORM_BATCH;  // a multiple of 4

// -------------------------------------------------------------- density --
// computeDensities (Algorithm.cu:17-54) and computeBoundaryDensityContribution
// (BoundaryAlgorithm.cu:15-51). Phase 1 finds each particle's accepted
// neighbours in the reference's visiting order and stores their slots in the
// neighbour list; phase 2 accumulates the density over that list in the same
// order with every lane active.
//
// List entries travel in chunks of kChunk = 4 (one 16-byte segment, written
// whole from an LDS staging ring, read with one 16-byte load); the layout of
// the chunks is listChunk's.
constexpr uint32_t kChunk = 4;

// The consumers stream the list through once each; non-temporal loads keep it
// from evicting the particle records that they gather (the list stores stay
// ordinary: non-temporal stores measured 3x slower in the list builder).
typedef uint32_t U32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint4 loadChunk(const uint32_t* p)
{
    const U32x4 v = __builtin_nontemporal_load(reinterpret_cast<const U32x4*>(p));
    return uint4 {v.x, v.y, v.z, v.w};
}

__device__ __forceinline__ void storeChunk(uint4* p, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    *p = uint4 {a, b, c, d};
}

// Wave-blocked list layout: the 64 slots of an aligned wave (t & ~63) own
// listCap * 64 consecutive entries, row by row; a row holds kRowEntries
// entries of each of the 64 slots (kRowChunks 16-byte chunks per slot, side
// by side), so chunk k / 4 of slot t is the 16 bytes at chunk index
// ((t / 64) * (listCap / kRowEntries) + k / kRowEntries) * 64 * kRowChunks
// + (t % 64) * kRowChunks + (k / 4) % kRowChunks. A wave reads or writes its
// rows contiguously, and its whole list is one contiguous block (few pages
// per wave, whatever the pool).
constexpr uint32_t kRowChunks = kListRowChunks;
constexpr uint32_t kRowEntries = kListRowEntries;
__device__ __forceinline__ size_t listChunk(const Ctx& c, uint32_t t, uint32_t k)
{
    return (static_cast<size_t>(t >> 6) * (c.listCap / kRowEntries) + k / kRowEntries) * (64U * kRowChunks) +
           (t & 63U) * kRowChunks + (k / kChunk) % kRowChunks;
}

__device__ __forceinline__ const uint32_t* listAt(const Ctx& c, uint32_t t, uint32_t k)
{
    return c.list + listChunk(c, t, k) * kChunk + (k % kChunk);
}

// Chunk k / 4 of a list whose first chunk is `row0` (listChunk(c, t, 0)):
// rows lie 64 * kRowChunks chunks apart.
__device__ __forceinline__ const uint32_t* listRowAt(const uint4* row0, uint32_t k)
{
    return reinterpret_cast<const uint32_t*>(row0 + (k / kRowEntries) * (64U * kRowChunks) + (k / kChunk) % kRowChunks);
}

// The T at the start of 16-byte record j of a particle array, addressed by a
// 32-bit byte offset from the array base (a scalar base plus a 32-bit vector
// offset: one shift per gather instead of 64-bit address arithmetic). The
// engine bounds the particle pool below 2^28 records, so j * 16 fits.
template <class T>
__device__ __forceinline__ const T& recAt(const F4* base, uint32_t j)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (j << 4U));
}

__device__ __forceinline__ float waveMinF(float v)
{
    for (int off = 32; off > 0; off >>= 1)
    {
        v = fminf(v, __shfl_xor(v, off, 64));
    }
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

__device__ __forceinline__ float waveMaxF(float v)
{
    for (int off = 32; off > 0; off >>= 1)
    {
        v = fmaxf(v, __shfl_xor(v, off, 64));
    }
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// Wave-uniform walk with a lane-parallel candidate filter (the list builder's
// walk). The 64 lanes of a wave (all of one scene) visit the union of their
// candidate cells in the reference's order (x, then y, then z innermost); a
// lane accepts a candidate only if its cell lies in the lane's own candidate
// box and d^2 <= (2 h_j)^2. A cell's slots are taken 64 at a time,
// one per lane, and each is tested against the box of the wave's queries: a
// candidate farther from that box than its own support 2h_j cannot be
// accepted by any lane. The survivors are packed in ascending slot order into
// the wave's LDS table (`wstash`, `wslot`), and every lane applies the exact
// test to each of them, so each lane still sees exactly its reference
// neighbour sequence. push(hit, j) is called for every survivor (hit: the
// lane accepts slot j), drain() after every four.
//
// The exact test needs no per-survivor conditions besides d^2 <= (2 h_j)^2:
// the table entries past the S survivors of a batch (read by its last round
// of four) carry a support of -1, which no distance passes, and a lane whose
// own box misses the cell tests with a NaN x coordinate, whose d^2 is NaN and
// fails the comparison.
//
// The box distance is a lower bound of every lane's distance in FP32 as well:
// each rounded difference, square and sum is monotonic in its operands and
// uses the same association as the exact test; the margin is extra.
constexpr uint32_t kStashPad = 4;  // table entries past 64: sentinels read by the last round

template <class P, class D>
__device__ __forceinline__ void waveWalkFiltered(const Ctx& c, const SceneDev& sc, uint32_t lost, bool valid,
                                                 const CellBox& b, const F4& q, F4* wstash, uint32_t* wslot,
                                                 uint32_t lane, P&& push, D&& drain, uint32_t& survivors)
{
    const int ux0 = waveMinI(valid ? b.x0 : 0x7FFFFFFF);
    const int uy0 = waveMinI(valid ? b.y0 : 0x7FFFFFFF);
    const int uz0 = waveMinI(valid ? b.z0 : 0x7FFFFFFF);
    const int ux1 = waveMaxI(valid ? b.x1 : -1);
    const int uy1 = waveMaxI(valid ? b.y1 : -1);
    const int uz1 = waveMaxI(valid ? b.z1 : -1);
    const float qx0 = waveMinF(valid ? q.x : FLT_MAX);
    const float qy0 = waveMinF(valid ? q.y : FLT_MAX);
    const float qz0 = waveMinF(valid ? q.z : FLT_MAX);
    const float qx1 = waveMaxF(valid ? q.x : -FLT_MAX);
    const float qy1 = waveMaxF(valid ? q.y : -FLT_MAX);
    const float qz1 = waveMaxF(valid ? q.z : -FLT_MAX);
    const int gx = sc.gx;
    const int gxgy = sc.gx * sc.gy;
    const uint32_t base = sc.cellBase;
    // the union box's cells in visiting order (z fastest, then y, then x),
    // 64 at a time: lane l loads the slot range of the group's cell l in one
    // vector load, and the walk takes each cell's range from its lane, so no
    // cell waits on a load of its own range
    const uint32_t nz = static_cast<uint32_t>(uz1 - uz0 + 1);
    const uint32_t nyz = static_cast<uint32_t>(uy1 - uy0 + 1) * nz;
    const uint32_t ncell = static_cast<uint32_t>(ux1 - ux0 + 1) * nyz;
    int x = ux0;
    int y = uy0;
    int z = uz0;
    for (uint32_t k0 = 0; k0 < ncell; k0 += 64U)
    {
        uint32_t vlo = 0;
        uint32_t vhi = 0;
        const uint32_t kl = k0 + lane;
        if (kl < ncell)
        {
            const uint32_t lx = kl / nyz;
            const uint32_t ly = (kl - lx * nyz) / nz;
            const uint32_t lz = kl - lx * nyz - ly * nz;
            const uint32_t cl = base + static_cast<uint32_t>((ux0 + static_cast<int>(lx)) +
                                                             (uy0 + static_cast<int>(ly)) * gx +
                                                             (uz0 + static_cast<int>(lz)) * gxgy);
            vlo = max(c.cellOff[cl], lost);
            vhi = c.cellOff[cl + 1U];
        }
        const uint32_t kn = min(ncell - k0, 64U);
        for (uint32_t k = 0; k < kn; ++k)
        {
            const bool inbox = valid && x >= b.x0 && x <= b.x1 && y >= b.y0 && y <= b.y1 && z >= b.z0 && z <= b.z1;
            if (++z > uz1)
            {
                z = uz0;
                if (++y > uy1)
                {
                    y = uy0;
                    ++x;
                }
            }
            if (__ballot(inbox ? 1 : 0) == 0ULL)
            {
                continue;
            }
            {
                const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(vlo), static_cast<int>(k)));
                const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(vhi), static_cast<int>(k)));
                const float qxc = inbox ? q.x : __builtin_nanf("");
                for (uint32_t j0 = lo; j0 < hi; j0 += 64U)
                {
                    const uint32_t j = j0 + lane;
                    const bool in = j < hi;
                    const F4 o = c.s0[in ? j : hi - 1U];
                    const float ex = fmaxf(fmaxf(qx0 - o.x, o.x - qx1), 0.0F);
                    const float ey = fmaxf(fmaxf(qy0 - o.y, o.y - qy1), 0.0F);
                    const float ez = fmaxf(fmaxf(qz0 - o.z, o.z - qz1), 0.0F);
                    const float d2b = (ex * ex + ey * ey) + ez * ez;
                    const bool keep = in && d2b <= o.w * 1.0001F;
                    const uint64_t m = __ballot(keep ? 1 : 0);
                    if (m == 0ULL)
                    {
                        continue;
                    }
                    const uint32_t S = static_cast<uint32_t>(__builtin_popcountll(m));
                    survivors += S;
                    if (keep)
                    {
                        // rank among the survivors = survivors in lower lanes
                        const uint32_t r = __builtin_amdgcn_mbcnt_hi(
                            static_cast<uint32_t>(m >> 32U), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0U));
                        wstash[r] = o;
                        wslot[r] = j;
                    }
                    if (lane < 4U)
                    {
                        wstash[S + lane].w = -1.0F;  // sentinels of the last round
                    }
                    for (uint32_t r = 0; r < S; r += 4U)
                    {
                        F4 oc[4];
                        uint32_t jj[4];
#pragma unroll
                        for (uint32_t u = 0; u < 4U; ++u)
                        {
                            oc[u] = wstash[r + u];
                            jj[u] = wslot[r + u];
                        }
#pragma unroll
                        for (uint32_t u = 0; u < 4U; ++u)
                        {
                            const float dx = oc[u].x - qxc;
                            const float dy = oc[u].y - q.y;
                            const float dz = oc[u].z - q.z;
                            const float d2 = dot3w(dx, dy, dz, dx, dy, dz);
                            // every lane tests, no branch
                            push(d2 <= oc[u].w, jj[u]);
                        }
                        drain();
                    }
                }
            }
        }
    }
}

// Lane l's value of v (wave-uniform result).
__device__ __forceinline__ float laneValue(float v, uint32_t l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), static_cast<int>(l)));
}

// The walk of one particle by a whole wave (walkFluid's cells, candidates and
// order; q and b wave-uniform), for particles whose neighbours overflow their
// list: a cell's slots are taken 64 at a time, one per lane; each lane whose
// candidate is accepted (d^2 <= (2 h_j)^2) evaluates term(j, d2), then
// fold(m) runs on every lane with the mask m of the accepting lanes, which it
// folds in ascending lane order, the visiting order. A sum folded as
// acc = acc + (lane l's term) is the serial walk's sum bit for bit; the
// terms, the expensive part, are evaluated 64 at a time.
template <class T, class F>
__device__ __forceinline__ void waveWalkFold(const Ctx& c, const SceneDev& sc, uint32_t lost, const CellBox& b,
                                             const F4& q, uint32_t lane, T&& term, F&& fold)
{
    ConstU32* off = (ConstU32*)(c.cellOff);
    const int gxgy = sc.gx * sc.gy;
    for (int x = b.x0; x <= b.x1; ++x)
    {
        for (int y = b.y0; y <= b.y1; ++y)
        {
            for (int z = b.z0; z <= b.z1; ++z)
            {
                const uint32_t cell = sc.cellBase + static_cast<uint32_t>(x + y * sc.gx + z * gxgy);
                const uint32_t lo = __builtin_amdgcn_readfirstlane(max(off[cell], lost));
                const uint32_t hi = __builtin_amdgcn_readfirstlane(off[cell + 1U]);
                for (uint32_t j0 = lo; j0 < hi; j0 += 64U)
                {
                    const uint32_t j = j0 + lane;
                    const bool in = j < hi;
                    const F4 o = c.s0[in ? j : hi - 1U];
                    const float dx = o.x - q.x;
                    const float dy = o.y - q.y;
                    const float dz = o.z - q.z;
                    const float d2 = dot3w(dx, dy, dz, dx, dy, dz);
                    const bool acc = in && d2 <= o.w;
                    const uint64_t m = __ballot(acc ? 1 : 0);
                    if (m == 0ULL)
                    {
                        continue;
                    }
                    if (acc)
                    {
                        term(j, d2);
                    }
                    fold(m);
                }
            }
        }
    }
}

// computeBoundaryDensityContribution (BoundaryAlgorithm.cu:15-51) for one
// particle: sum of psi_b W(d, h_i) over boundary particles within 2 h_i.
__device__ __forceinline__ float boundaryDensity(const Ctx& c, const SceneDev& sc, const CellBox& b, const F4& q,
                                              float hi)
{
    const float R = 2.0F * hi;
    const float r2 = R * R;
    const int gxgy = sc.gx * sc.gy;
    float rhoB = 0.0F;
    for (int x = b.x0; x <= b.x1; ++x)
    {
        for (int y = b.y0; y <= b.y1; ++y)
        {
            for (int z = b.z0; z <= b.z1; ++z)
            {
                const uint32_t cell = sc.cellBase + static_cast<uint32_t>(x + y * sc.gx + z * gxgy);
                const uint32_t start = c.bcellOff[cell];
                const uint32_t lo = max(start, sc.bLostEnd);
                const uint32_t hiS = c.bcellOff[cell + 1U];
                if (lo >= hiS || !boxMayReach(c.bbox0[cell], c.bbox1[cell], q, r2))
                {
                    continue;
                }
                // chunks of the cell's range, each culled by its box
                for (uint32_t k0 = start + (lo - start) / kBoundaryGroup * kBoundaryGroup; k0 < hiS;
                     k0 += kBoundaryGroup)
                {
                    if (!boxMayReach(c.bgbox0[k0], c.bgbox1[k0], q, r2))
                    {
                        continue;
                    }
                    const uint32_t jhi = min(hiS, k0 + kBoundaryGroup);
                    for (uint32_t j = max(lo, k0); j < jhi; ++j)
                    {
                        const F4 o = c.bs0[j];
                        const float dx = o.x - q.x;
                        const float dy = o.y - q.y;
                        const float dz = o.z - q.z;
                        const float d2 = dot3w(dx, dy, dz, dx, dy, dz);
                        if (d2 <= r2)
                        {
                            rhoB = rhoB + o.w * wendland(sqrtf(d2), hi);
                        }
                    }
                }
            }
        }
    }
    return rhoB;
}

// Adds the boundary contribution (launched unconditionally by the
// reference), then Tait pressure; stores the force kernels' gather records
// fa = (position, p / rho^2) and fb = (velocity, rho).
__device__ __forceinline__ void finishDensity(const Ctx& c, const SceneDev& sc, const CellBox& b, const F4& q,
                                              float hi, uint32_t t, float rho)
{
    float rhoB = 0.0F;
    if (sc.bCount > 0U && touchesEdge(sc, b) && wallsMayReach(c, t, 2.0F * hi))
    {
        rhoB = boundaryDensity(c, sc, b, q, hi);
    }
    rho = rho + rhoB;
    const float p = tait(rho, sc);
    const F4 v = c.s1[t];
    c.fa[t] = F4 {q.x, q.y, q.z, p / (rho * rho)};
    c.fb[t] = F4 {v.x, v.y, v.z, rho};
}

// sqrtf(d2) / h0 through sqrtRanged and y0 = divRecip(h0), h0 in
// [2^-40, 2^39] (the uniform scenes' condition): for d2 = +0 or d2 in
// [2^-80, 2^80] the root is +0 or lies in [2^-40, 2^40], inside both
// functions' domains. A list entry of such a scene has d2 <= (2 h0)^2 <= 2^80
// (its acceptance test), so only d2 in (0, 2^-80) (rare) takes sqrtf and the
// division: one unsigned comparison of the bits (d2 is a sum of squares, +0 or
// positive; for +0 the subtraction wraps).
__device__ __forceinline__ float quotientH0(float d2, float h0, float y0)
{
    float q = divBy(sqrtRanged(d2), h0, y0);
    if (__float_as_uint(d2) - 1U < __float_as_uint(0x1p-80F) - 1U)
    {
        q = sqrtf(d2) / h0;
    }
    return q;
}

// Phase 2 of computeDensities over the neighbour list, plus
// computeBoundaryDensityContribution; also precomputes p / rho^2. Particles
// whose neighbours overflow the list are left to kDensityOverflow.
// computeDensities (Algorithm.cu:17-54) for sorted slot t over its
// neighbour list (cnt <= listCap entries), in list order: the reference's
// visiting order. UB and GB are the entries whose data are loaded together in
// the uniform-h/m and general loops.
template <uint32_t UB, uint32_t GB>
__device__ __forceinline__ float listDensity(const Ctx& c, const SceneDev& sc, uint32_t lp, uint32_t cnt, const F4& q)
{
    float rho = 0.0F;
    if ((sc.flags & kSceneUniformHM) != 0U)
    {
        // every neighbour has h0 and m0: the same sum with one gather
        // (the position) per pair; h0 lies in [2^-40, 2^39] (a condition of
        // the flag), so d / h0 goes through its reciprocal
        const float h0 = sc.h0;
        const float m0 = sc.m0;
        const float c3 = coefOverH3(h0);
        const float y0 = divRecip(h0);
        const uint4* row0 = reinterpret_cast<const uint4*>(c.list) + listChunk(c, lp, 0U);
        uint32_t e = 0;
        for (; e + UB <= cnt; e += UB)
        {
            uint32_t j[UB];
            float3 o[UB];
#pragma unroll
            for (uint32_t u = 0; u < UB; u += 4U)
            {
                const uint4 v = loadChunk(listRowAt(row0, e + u));
                j[u] = v.x;
                j[u + 1U] = v.y;
                j[u + 2U] = v.z;
                j[u + 3U] = v.w;
            }
#pragma unroll
            for (uint32_t u = 0; u < UB; ++u)
            {
                o[u] = recAt<float3>(c.s0, j[u]);
            }
#pragma unroll
            for (uint32_t u = 0; u < UB; ++u)
            {
                const float dx = o[u].x - q.x;
                const float dy = o[u].y - q.y;
                const float dz = o[u].z - q.z;
                rho = rho + m0 * wendlandQ(quotientH0(dot3w(dx, dy, dz, dx, dy, dz), h0, y0), c3);
            }
        }
        for (; e < cnt; ++e)
        {
            const float3 o = recAt<float3>(c.s0, listRowAt(row0, e)[e % kChunk]);
            const float dx = o.x - q.x;
            const float dy = o.y - q.y;
            const float dz = o.z - q.z;
            rho = rho + m0 * wendlandQ(quotientH0(dot3w(dx, dy, dz, dx, dy, dz), h0, y0), c3);
        }
    }
    else
    {
        const uint4* row0 = reinterpret_cast<const uint4*>(c.list) + listChunk(c, lp, 0U);
        uint32_t e = 0;
        for (; e + GB <= cnt; e += GB)
        {
            uint32_t j[GB];
            F4 o[GB];
            F4 hm[GB];
#pragma unroll
            for (uint32_t u = 0; u < GB; u += 4U)
            {
                const uint4 v = loadChunk(listRowAt(row0, e + u));
                j[u] = v.x;
                j[u + 1U] = v.y;
                j[u + 2U] = v.z;
                j[u + 3U] = v.w;
            }
#pragma unroll
            for (uint32_t u = 0; u < GB; ++u)
            {
                o[u] = recAt<F4>(c.s0, j[u]);
                hm[u] = recAt<F4>(c.shm, j[u]);
            }
#pragma unroll
            for (uint32_t u = 0; u < GB; ++u)
            {
                const float dx = o[u].x - q.x;
                const float dy = o[u].y - q.y;
                const float dz = o[u].z - q.z;
                const float d = sqrtf(dot3w(dx, dy, dz, dx, dy, dz));
                rho = rho + hm[u].y * wendlandC(d, hm[u].x, hm[u].z);
            }
        }
        for (; e < cnt; ++e)
        {
            const uint32_t jj = listRowAt(row0, e)[e % kChunk];
            const F4 o = recAt<F4>(c.s0, jj);
            const F4 hm = recAt<F4>(c.shm, jj);
            const float dx = o.x - q.x;
            const float dy = o.y - q.y;
            const float dz = o.z - q.z;
            rho = rho + hm.y * wendlandC(sqrtf(dot3w(dx, dy, dz, dx, dy, dz)), hm.x, hm.z);
        }
    }
    return rho;
}

// ---- tolerance mode (kSceneToleranceMath) ----------------------------------
// The same sums in the same list order, with d from the hardware reciprocal
// square root (d = d2 * rsq(d2)) and 1 / h_j from the hardware reciprocal:
// each term within a few ulps of the reference's, the sums within 1e-5 of
// them. List entries satisfy d2 <= (2 h_j)^2 (q <= 2 up to rounding, where
// the kernel is 0 to within 1e-29), so the kernels need no range branch; the
// particle itself (d2 = 0) is evaluated at d2 = kTinyD2, which gives W(0) and
// a zero gradient term.
constexpr float kTinyD2 = 1e-30F;

__device__ __forceinline__ float fastDistOverH(float d2, float invh)
{
    const float dd = fmaxf(d2, kTinyD2);
    return (dd * __builtin_amdgcn_rsqf(dd)) * invh;
}

// (1 - q/2)^4 (2q + 1): the Wendland kernel over its per-particle factor
__device__ __forceinline__ float fastWendlandShape(float qh)
{
    const float t = __builtin_fmaf(-0.5F, qh, 1.0F);
    const float t2 = t * t;
    return (t2 * t2) * __builtin_fmaf(2.0F, qh, 1.0F);
}

template <uint32_t UB>
__device__ __forceinline__ float listDensityFast(const Ctx& c, const SceneDev& sc, uint32_t lp, uint32_t cnt,
                                                 const F4& q)
{
    float rho = 0.0F;
    const uint4* row0 = reinterpret_cast<const uint4*>(c.list) + listChunk(c, lp, 0U);
    if ((sc.flags & kSceneUniformHM) != 0U)
    {
        const float invh = 1.0F / sc.h0;
        const float mc3 = sc.m0 * coefOverH3(sc.h0);
        uint32_t e = 0;
        for (; e + UB <= cnt; e += UB)
        {
            uint32_t j[UB];
            float3 o[UB];
#pragma unroll
            for (uint32_t u = 0; u < UB; u += 4U)
            {
                const uint4 v = loadChunk(listRowAt(row0, e + u));
                j[u] = v.x;
                j[u + 1U] = v.y;
                j[u + 2U] = v.z;
                j[u + 3U] = v.w;
            }
#pragma unroll
            for (uint32_t u = 0; u < UB; ++u)
            {
                o[u] = recAt<float3>(c.s0, j[u]);
            }
#pragma unroll
            for (uint32_t u = 0; u < UB; ++u)
            {
                const float dx = o[u].x - q.x;
                const float dy = o[u].y - q.y;
                const float dz = o[u].z - q.z;
                const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                rho = __builtin_fmaf(mc3, fastWendlandShape(fastDistOverH(d2, invh)), rho);
            }
        }
        for (; e < cnt; ++e)
        {
            const float3 o = recAt<float3>(c.s0, listRowAt(row0, e)[e % kChunk]);
            const float dx = o.x - q.x;
            const float dy = o.y - q.y;
            const float dz = o.z - q.z;
            const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            rho = __builtin_fmaf(mc3, fastWendlandShape(fastDistOverH(d2, invh)), rho);
        }
    }
    else
    {
        for (uint32_t e = 0; e < cnt; e += 4U)
        {
            const uint4 v = loadChunk(listRowAt(row0, e));
            const uint32_t j[4] = {v.x, v.y, v.z, v.w};
            F4 o[4];
            F4 hm[4];
#pragma unroll
            for (uint32_t u = 0; u < 4U; ++u)
            {
                const uint32_t jj = e + u < cnt ? j[u] : j[0];
                o[u] = recAt<F4>(c.s0, jj);
                hm[u] = recAt<F4>(c.shm, jj);
            }
#pragma unroll
            for (uint32_t u = 0; u < 4U; ++u)
            {
                const float dx = o[u].x - q.x;
                const float dy = o[u].y - q.y;
                const float dz = o[u].z - q.z;
                const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                const float w = hm[u].y * (hm[u].z * fastWendlandShape(fastDistOverH(d2, __builtin_amdgcn_rcpf(hm[u].x))));
                rho = rho + (e + u < cnt ? w : 0.0F);
            }
        }
    }
    return rho;
}

// Neighbour lists and densities (computeDensities, Algorithm.cu:17-54): each
// thread builds its particle's list, then sums the density over it. BS (the
// block size) is a compile-time constant so that the staging addresses fold
// into immediates.
// Most cells in a wave's union box for the wave-uniform walk. A compact wave
// of 64 queries spans 2-3 cells per axis (27-54 cells); a wave of the sparse
// layer that forms under a falling block or along a free surface (about one
// particle per cell along a row) spans a few hundred mostly empty cells and
// still walks faster as a wave than lane by lane, whose 27-cell chains of
// dependent loads make such a wave the kernel's last to finish; only
// scattered spray exceeds this. (1M dam break, ms/step over 1,000 steps, by
// limit: 64: 1.080, 256: 1.076, 384: 1.047, 512: 1.051, 1024: 1.060,
// unlimited: 1.131.)
#ifndef ASPH_MAX_UNION_CELLS
#define ASPH_MAX_UNION_CELLS 384
#endif
constexpr uint32_t kMaxUnionCells = ASPH_MAX_UNION_CELLS;

template <uint32_t BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(ASPH_NEIGHBOR_WAVES))) void kNeighbors(Ctx c)
{
    // staging ring of kRing entries per thread, entry-major ([e][thread]) so
    // that the lanes' writes fall in distinct banks; per wave, the survivor
    // table of the candidate filter
    constexpr uint32_t kRing = 2U * kRowEntries;
    constexpr uint32_t kTable = 64U + kStashPad;
    __shared__ uint32_t stage[kRing * BS];
    __shared__ F4 stash[(BS / 64U) * kTable];
    __shared__ uint32_t slots[(BS / 64U) * kTable];
    const uint32_t n = *c.nTotal;
    const uint32_t nPad = (n + 63U) & ~63U;
    const uint32_t K = c.listCap;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63U;
    F4* wstash = stash + (tid / 64U) * kTable;
    uint32_t* wslot = slots + (tid / 64U) * kTable;
    // one particle per thread (the grid covers the slots; a loop here would
    // let the compiler hoist the pressure's constants across the walk). The
    // grid's first kHeavyCap waves take the listed heavy waves (kGather),
    // the others the waves of 64 query positions in order, skipping listed
    // ones.
    uint32_t blk = blockIdx.x;
    if constexpr (kNeighborRunWaves != 0U)
    {
        // regular workgroups (past the heavy ones; kHeavyCap / (BS / 64) is a
        // multiple of 8, so a workgroup keeps its XCD) in runs of G
        // consecutive workgroups per XCD, the runs dealt to the XCDs in turn;
        // the tail past the last whole round of runs keeps its order
        constexpr uint32_t kHeavyBlocks = kHeavyCap / (BS / 64U);
        constexpr uint32_t G = kNeighborRunWaves / (BS / 64U) > 0U ? kNeighborRunWaves / (BS / 64U) : 1U;
        if (blk >= kHeavyBlocks)
        {
            const uint32_t rb = blk - kHeavyBlocks;
            const uint32_t needed = ((n + 63U) / 64U + BS / 64U - 1U) / (BS / 64U);
            const uint32_t avail = gridDim.x - kHeavyBlocks;
            const uint32_t used = needed < avail ? needed : avail;
            if (rb < (used / (8U * G)) * (8U * G))
            {
                const uint32_t k = rb / 8U;
                blk = kHeavyBlocks + (k / G) * (8U * G) + (rb % 8U) * G + k % G;
            }
        }
    }
    const uint32_t v = blk * (BS / 64U) + tid / 64U;
    uint32_t w;
    if (v < kHeavyCap)
    {
        if (v >= min(*c.heavyCount, kHeavyCap))
        {
            return;
        }
        w = c.heavyList[v];
    }
    else
    {
        w = v - kHeavyCap;
        if (w * 64U < nPad && c.wheavy[w] != 0U)
        {
            return;
        }
    }
    const uint32_t tw = w * 64U + lane;
    if (tw < nPad)
    {
        const bool valid = tw < n;
        // query position pq handles sorted slot t; the list sits at pq
        const uint32_t pq = valid ? tw : n - 1U;
        const uint32_t t = querySlot(c, pq);
        const uint32_t s = c.sscene[t];
        const uint32_t sw = __builtin_amdgcn_readfirstlane(s);
        const bool oneScene = __ballot((valid && s != sw) ? 1 : 0) == 0ULL;
        const SceneDev& sc = c.scenes[s];
        const uint32_t lost = c.states[s].lostEnd;
        const F4 q = c.s0[t];
        const float hi = c.shm[t].x;
        const float R = 2.0F * hi;
        const CellBox b = cellBox(sc, q.x, q.y, q.z, R);
        // Entries [done, cnt) are staged at ring positions (done .. cnt-1) mod
        // kRing; at most kRing - 1 are staged, so position cnt mod kRing is
        // free. The counters are kept as byte offsets into the ring: entry k
        // of the thread sits at (k mod kRing) * kRow + 4 tid, kRow = 4 BS, and
        // cntA = cnt * kRow + 4 tid (mod 2^32, a multiple of the ring's size),
        // so the next entry's offset is one mask and a hit one add. `done`
        // is also kept as a plain count, which gives cnt at the end.
        constexpr uint32_t kRow = 4U * BS;
        constexpr uint32_t kRingMask = kRing * kRow - 1U;
        uint32_t cntA = 4U * tid;
        uint32_t doneA = 4U * tid;
        uint32_t done = 0;
        char* ring = reinterpret_cast<char*>(stage);
        auto staged = [&](uint32_t off) { return *reinterpret_cast<const uint32_t*>(ring + off); };
        // the thread's next row (listChunk): rows of a wave's block are
        // 64 * kRowChunks chunks apart
        uint4* cptr = reinterpret_cast<uint4*>(c.list) + listChunk(c, pq, 0U);
        constexpr uint32_t cstep = 64U * kRowChunks;
        auto push = [&](bool hit, uint32_t j) {
            *reinterpret_cast<uint32_t*>(ring + (cntA & kRingMask)) = j;
            cntA += hit ? kRow : 0U;
        };
        // writes the oldest staged row once it is full (called after at most
        // kChunk pushes, so the ring never holds more than kRing - 1 entries);
        // the row's chunks leave together, so that the sectors they fill are
        // written whole
        auto drain = [&]() {
            if (cntA - doneA >= kRowEntries * kRow)
            {
                if (done < K)
                {
                    // done is a multiple of kRowEntries: the row does not wrap
#pragma unroll
                    for (uint32_t h = 0; h < kRowChunks; ++h)
                    {
                        const uint32_t e0 = (doneA + h * kChunk * kRow) & kRingMask;
                        storeChunk(cptr + h, staged(e0), staged(e0 + kRow), staged(e0 + 2U * kRow),
                                   staged(e0 + 3U * kRow));
                    }
                    cptr += cstep;
                }
                done += kRowEntries;
                doneA += kRowEntries * kRow;
            }
        };
        // A wave whose queries lie far apart (spray, a sparse surface) would
        // walk every cell of its union box: its lanes walk their own cells
        // instead, as do lanes of waves that straddle scenes.
        const int ux0 = waveMinI(valid ? b.x0 : 0x7FFFFFFF);
        const int uy0 = waveMinI(valid ? b.y0 : 0x7FFFFFFF);
        const int uz0 = waveMinI(valid ? b.z0 : 0x7FFFFFFF);
        const int ux1 = waveMaxI(valid ? b.x1 : -1);
        const int uy1 = waveMaxI(valid ? b.y1 : -1);
        const int uz1 = waveMaxI(valid ? b.z1 : -1);
        const float unionCells = static_cast<float>(ux1 - ux0 + 1) * static_cast<float>(uy1 - uy0 + 1) *
                                 static_cast<float>(uz1 - uz0 + 1);
        const bool waveWalk = oneScene && unionCells <= static_cast<float>(kMaxUnionCells);
        uint32_t survivors = 0;
        if (waveWalk)
        {
            waveWalkFiltered(c, c.scenes[sw], c.states[sw].lostEnd, valid, b, q, wstash, wslot, lane, push, drain,
                             survivors);
        }
        else if (valid)
        {
            walkFluid(c, sc, lost, b, q, [&](uint32_t j, const F4&, float) {
                push(true, j);
                drain();
            });
        }
        const uint32_t cnt = done + (cntA - doneA) / kRow;
        if (lane == 0U)
        {
            c.wcost[w] = waveWalk ? survivors : kCostPerLane;
            if (waveWalk && w % kCostSample == 0U)
            {
                atomicAdd(&c.wstat[0], survivors);
                atomicAdd(&c.wstat[1], 1U);
            }
        }
        if (waveWalk && c.queryTileBits == kQueryAuto && w % kQuerySample == 0U && c.states[sw].qorder == 0U)
        {
            // the automatic query order's measure, on one wave in kQuerySample
            // (atomics of every wave on one address would serialise):
            // survivors per wave against neighbours per particle
            uint32_t hits = valid ? cnt : 0U;
            for (int off = 32; off > 0; off >>= 1)
            {
                hits += static_cast<uint32_t>(__shfl_xor(static_cast<int>(hits), off, 64));
            }
            if (lane == 0U)
            {
                atomicAdd(&c.states[sw].qsurv, survivors);
                atomicAdd(&c.states[sw].qhits, hits);
            }
        }
        if (valid)
        {
            if (cnt <= K)
            {
                // partial last row (entries past cnt are never read)
#pragma unroll
                for (uint32_t h = 0; h < kRowChunks; ++h)
                {
                    if (cnt > done + h * kChunk)
                    {
                        const uint32_t e0 = (doneA + h * kChunk * kRow) & kRingMask;
                        storeChunk(cptr + h, staged(e0), staged(e0 + kRow), staged(e0 + 2U * kRow),
                                   staged(e0 + 3U * kRow));
                    }
                }
            }
            c.ncount[t] = cnt;
            if (cnt > K)
            {
                c.overflowList[atomicAdd(c.overflowCount, 1U)] = t;  // kDensityOverflow
            }
            else
            {
                // the density over the list just written (its entries and the
                // neighbours' records are still in cache), then the boundary
                // term, the pressure and the force records
                const float rho = (sc.flags & kSceneToleranceMath) != 0U
                                      ? listDensityFast<kListDensityUniformBatch>(c, sc, pq, cnt, q)
                                      : listDensity<kListDensityUniformBatch, kListDensityBatch>(c, sc, pq, cnt, q);
                finishDensity(c, sc, b, q, hi, t, rho);
            }
        }
    }
}

// Density of the particles whose neighbours overflow their list: the walk
// again (computeDensities, Algorithm.cu:17-54), one wave per particle,
// accumulating in the same order.
__global__ __launch_bounds__(256) void kDensityOverflow(Ctx c)
{
    const uint32_t m = *c.overflowCount;
    const uint32_t lane = threadIdx.x & 63U;
    const uint32_t waves = gridDim.x * (blockDim.x / 64U);
    for (uint32_t i = blockIdx.x * (blockDim.x / 64U) + threadIdx.x / 64U; i < m; i += waves)
    {
        const uint32_t t = c.overflowList[i];
        const uint32_t s = c.sscene[t];
        const SceneDev& sc = c.scenes[s];
        const F4 q = c.s0[t];
        const float hi = c.shm[t].x;
        const CellBox b = cellBox(sc, q.x, q.y, q.z, 2.0F * hi);
        float rho = 0.0F;
        float w = 0.0F;
        waveWalkFold(
            c, sc, c.states[s].lostEnd, b, q, lane,
            [&](uint32_t j, float d2) {
                const F4 hm = c.shm[j];
                w = hm.y * wendlandC(sqrtf(d2), hm.x, hm.z);
            },
            [&](uint64_t m) {
                while (m != 0ULL)
                {
                    rho = rho + laneValue(w, static_cast<uint32_t>(__builtin_ctzll(m)));
                    m &= m - 1ULL;
                }
            });
        if (lane == 0U)
        {
            finishDensity(c, sc, b, q, hi, t, rho);
        }
    }
}

// --------------------------------------------------------------- forces --
struct PairAcc
{
    float apx, apy, apz;
    float avx, avy, avz;
    float asx, asy, asz;
    float amx, amy, amz;  // Morris laminar term (extension)
};

// One fluid pair of computePressureAccelerations (Algorithm.cu:72-105),
// computeViscosityAccelerations (:129-171) and
// computeSurfaceTensionAccelerations (:192-218), accumulated separately.
// fa = (position, p/rho^2) and fb = (velocity, rho) of neighbour j; mj, hj,
// wdj = its mass, smoothing length and kWDCoef / h^4.
// With kSharedH, yh = divRecip(hj) and hj lies in [2^-40, 2^40], so d / hj
// is taken through the shared reciprocal (same bits as the division).
template <bool kSharedH = false, bool kLam = false, bool kViscNoST = false>
__device__ __forceinline__ void forcePair(const F4& fa, const F4& fb, float mj, float hj, float wdj, const F4& q,
                                          const F4& vi, float rhoi, float prri, float nuNum, float epsH2, bool doVisc,
                                          bool doST, float nuLam, PairAcc& A, float yh = 0.0F, bool use = true)
{
    // use == false (the particle itself) computes and discards: the pair is
    // evaluated unconditionally so that the callers' gathers are not sunk
    // into a branch (which serialises their latency). kViscNoST: viscosity
    // on and surface tension off, known at compile time (no branch on them).
    const bool visc = kViscNoST ? true : doVisc;
    const bool st = kViscNoST ? false : doST;
    const F4& o = fa;
    const float ox = q.x - o.x;
    const float oy = q.y - o.y;
    const float oz = q.z - o.z;
    const float d2 = dot3w(ox, oy, oz, ox, oy, oz);
    float d;
    float dx = 0.0F;
    float dy = 1.0F;
    float dz = 0.0F;
    float qh;
    // every component in [2^-40, 2^39] puts d2 in [2^-80, 2^80] and d in
    // [2^-40, 2^40]: then the root goes through sqrtRanged and the four
    // quotients (three over d, one over hj) through reciprocals
    const float lo = fminf(fminf(fabsf(ox), fabsf(oy)), fabsf(oz));
    const float hi = fmaxf(fmaxf(fabsf(ox), fabsf(oy)), fabsf(oz));
    bool fast = (lo >= kDivLo && hi <= 0x1p39F) || !use;
    if (!fast)
    {
        // a component of exactly +-0 (particles on one lattice line): its
        // quotient through the reciprocal is a zero whose sign may differ
        // from the division's, and a zero of either sign leaves the
        // accumulators unchanged (they start at +0, and neither a - (+-0) nor
        // a + (+-0) ever yields -0 from them), so such a pair takes the fast
        // path as well when its other components are in range and d2 > 0
        const bool okx = ox == 0.0F || fabsf(ox) >= kDivLo;
        const bool oky = oy == 0.0F || fabsf(oy) >= kDivLo;
        const bool okz = oz == 0.0F || fabsf(oz) >= kDivLo;
        fast = okx && oky && okz && hi <= 0x1p39F && d2 >= 0x1p-80F;
    }
    // the reciprocal path's results on every lane (the particle itself gets
    // NaN quotients, which the selects below discard); lanes outside its
    // domain (rare) recompute them with the divisions
    d = sqrtRanged(d2);
    const float yd = divRecip(d);
    dx = divBy(ox, d, yd);
    dy = divBy(oy, d, yd);
    dz = divBy(oz, d, yd);
    qh = kSharedH ? divBy(d, hj, yh) : d / hj;
    if (!fast)
    {
        d = sqrtf(d2);
        dx = 0.0F;
        dy = 1.0F;
        dz = 0.0F;
        if (d > 0.0F)
        {
            dx = ox / d;
            dy = oy / d;
            dz = oz / d;
        }
        qh = d / hj;
    }
    const float pterm = prri + fa.w;
    const float dw = wendlandGradQ(qh, wdj);
    const float px = A.apx - ((mj * dx) * dw) * pterm;
    const float py = A.apy - ((mj * dy) * dw) * pterm;
    const float pz = A.apz - ((mj * dz) * dw) * pterm;
    A.apx = use ? px : A.apx;
    A.apy = use ? py : A.apy;
    A.apz = use ? pz : A.apz;
    if (visc)
    {
        const float vx = vi.x - fb.x;
        const float vy = vi.y - fb.y;
        const float vz = vi.z - fb.z;
        const float comp = dot3w(vx, vy, vz, ox, oy, oz);
        if (use && comp < 0.0F)
        {
            const float nu = nuNum / (rhoi + fb.w);
            const float piv = ((-nu) * comp) / (d2 + epsH2);
            const float sv = (mj * piv) * dw;
            A.avx = A.avx - sv * dx;
            A.avy = A.avy - sv * dy;
            A.avz = A.avz - sv * dz;
        }
    }
    if (kLam && use)
    {
        // Morris laminar term (extension): m_j nu (rho_i + rho_j) / (rho_i rho_j)
        // * (r . grad W) / (r^2 + 0.01 h_i^2) * (v_i - v_j), r . grad W = d W'
        const float k = ((mj * nuLam) * (rhoi + fb.w)) / (rhoi * fb.w);
        const float sl = (k * (d * dw)) / (d2 + epsH2);
        A.amx = A.amx + sl * (vi.x - fb.x);
        A.amy = A.amy + sl * (vi.y - fb.y);
        A.amz = A.amz + sl * (vi.z - fb.z);
    }
    if (st && use)
    {
        const float sw = mj * wendland(d, hj);
        A.asx = A.asx + sw * dx;
        A.asy = A.asy + sw * dy;
        A.asz = A.asz + sw * dz;
    }
}

struct BoundaryAcc
{
    float bpx, bpy, bpz;
    float bfx, bfy, bfz;
};

// computeBoundaryPressureAcceleration and computeBoundaryFrictionAcceleration
// (BoundaryAlgorithm.cu:53-162) for one particle; ri = (rho, p / rho^2, h,
// kWDCoef / h^4) of the particle.
__device__ __forceinline__ BoundaryAcc boundaryForce(const Ctx& c, const SceneDev& sc, const CellBox& b, const F4& q,
                                                  const F4& vi, const F4& ri, float epsH2)
{
    const float hi = ri.z;
    const float rhoi = ri.x;
    const float prri = ri.y;
    const float R = 2.0F * hi;
    const float r2 = R * R;
    const int gxgy = sc.gx * sc.gy;
    BoundaryAcc B {0.F, 0.F, 0.F, 0.F, 0.F, 0.F};
    for (int x = b.x0; x <= b.x1; ++x)
    {
        for (int y = b.y0; y <= b.y1; ++y)
        {
            for (int z = b.z0; z <= b.z1; ++z)
            {
                const uint32_t cell = sc.cellBase + static_cast<uint32_t>(x + y * sc.gx + z * gxgy);
                const uint32_t start = c.bcellOff[cell];
                const uint32_t lo = max(start, sc.bLostEnd);
                const uint32_t hiS = c.bcellOff[cell + 1U];
                if (lo >= hiS || !boxMayReach(c.bbox0[cell], c.bbox1[cell], q, r2))
                {
                    continue;
                }
                // chunks of the cell's range, each culled by its box
                for (uint32_t k0 = start + (lo - start) / kBoundaryGroup * kBoundaryGroup; k0 < hiS;
                     k0 += kBoundaryGroup)
                {
                    if (!boxMayReach(c.bgbox0[k0], c.bgbox1[k0], q, r2))
                    {
                        continue;
                    }
                    const uint32_t jhi = min(hiS, k0 + kBoundaryGroup);
                    for (uint32_t j = max(lo, k0); j < jhi; ++j)
                    {
                        const F4 o = c.bs0[j];
                        const float ox = q.x - o.x;
                        const float oy = q.y - o.y;
                        const float oz = q.z - o.z;
                        const float d2 = dot3w(ox, oy, oz, ox, oy, oz);
                        if (d2 > r2)
                        {
                            continue;
                        }
                        const float d = sqrtf(d2);
                        float dx = 0.0F;
                        float dy = 1.0F;
                        float dz = 0.0F;
                        if (d > 0.0F)
                        {
                            dx = ox / d;
                            dy = oy / d;
                            dz = oz / d;
                        }
                        const float psi = o.w;
                        const float dw = wendlandGradC(d, hi, ri.w);
                        const float sp = (psi * prri) * dw;
                        B.bpx = B.bpx - sp * dx;
                        B.bpy = B.bpy - sp * dy;
                        B.bpz = B.bpz - sp * dz;
                        const float comp = dot3w(vi.x, vi.y, vi.z, ox, oy, oz);
                        if (comp < 0.0F)
                        {
                            const float f = c.bvisc[j];
                            const float nu = ((f * hi) * sc.c0) / (2.0F * rhoi);
                            const float piv = ((-nu) * comp) / (d2 + epsH2);
                            B.bfx = B.bfx - ((dx * psi) * piv) * dw;
                            B.bfy = B.bfy - ((dy * psi) * piv) * dw;
                            B.bfz = B.bfz - ((dz * psi) * piv) * dw;
                        }
                    }
                }
            }
        }
    }
    return B;
}

// computeExternalAccelerations + fluid pair forces + boundary pressure and
// friction (BoundaryAlgorithm.cu:53-162), added to g in the reference's launch
// order (SphSimulation.cu:110-115), then the second half kick (:117) unless the
// scene refines this step (the kick then follows the refinement).
// g plus the fluid terms of A, in that order.
__device__ __forceinline__ F4 fluidAcceleration(const SceneDev& sc, const SceneState& st, const F4& vi, bool doVisc,
                                                bool doST, float nuLam, const PairAcc& A)
{
    float ax = st.gEff[0];
    float ay = st.gEff[1];
    float az = st.gEff[2];
    ax = ax + A.apx;
    ay = ay + A.apy;
    az = az + A.apz;
    if (doVisc)
    {
        ax = ax + A.avx;
        ay = ay + A.avy;
        az = az + A.avz;
    }
    if (nuLam != 0.0F)
    {
        ax = ax + A.amx;
        ay = ay + A.amy;
        az = az + A.amz;
    }
    if (doST)
    {
        const float k = sc.sigma / vi.w;
        ax = ax + k * A.asx;
        ay = ay + k * A.asy;
        az = az + k * A.asz;
    }
    return F4 {ax, ay, az, 0.0F};
}

// Stores the acceleration and density of sorted slot t, then the second half
// kick (halfKickVelocities, Algorithm.cu:270-284) unless the scene refines
// this step.
__device__ __forceinline__ void storeAndKick(const Ctx& c, const SceneDev& sc, const SceneState& st, const F4& vi,
                                             float rhoi, uint32_t t, float ax, float ay, float az)
{
    const uint32_t i = c.sortedVals[t];
    c.acc[i] = F4 {ax, ay, az, 0.0F};
    c.rho[i] = rhoi;
    if (st.refineNow == 0U)
    {
        const float hdt = sc.halfDt;
        float vx = vi.x + ax * hdt;
        float vy = vi.y + ay * hdt;
        float vz = vi.z + az * hdt;
        float vw = 0.0F;
        const float len = sqrtf(dot4(vx, vy, vz, vw));
        if (len > sc.vmax)
        {
            const float k = sc.vmax / len;
            vx = vx * k;
            vy = vy * k;
            vz = vz * k;
            vw = vw * k;
        }
        c.vel[i] = F4 {vx, vy, vz, vw};
    }
}

// The boundary pressure then friction sums added to a (SphSimulation.cu:115),
// then the store and kick.
__device__ __forceinline__ void addBoundaryAndFinish(const Ctx& c, const SceneDev& sc, const SceneState& st,
                                                     const CellBox& b, const F4& q, const F4& vi, const F4& ri,
                                                     float epsH2, uint32_t t, F4 a)
{
    const BoundaryAcc B = boundaryForce(c, sc, b, q, vi, ri, epsH2);
    float ax = a.x + B.bpx;
    float ay = a.y + B.bpy;
    float az = a.z + B.bpz;
    ax = ax + B.bfx;
    ay = ay + B.bfy;
    az = az + B.bfz;
    storeAndKick(c, sc, st, vi, ri.x, t, ax, ay, az);
}

__device__ __forceinline__ bool wallsInReach(const Ctx& c, const SceneDev& sc, const CellBox& b, uint32_t t, float hi)
{
    return sc.bCount > 0U && touchesEdge(sc, b) && wallsMayReach(c, t, 2.0F * hi);
}

// Everything after the fluid sums, in one place (the walking kernels).
__device__ __forceinline__ void finishForce(const Ctx& c, const SceneDev& sc, const SceneState& st, const CellBox& b,
                                            const F4& q, const F4& vi, const F4& ri, float epsH2, bool doVisc,
                                            bool doST, float nuLam, uint32_t t, const PairAcc& A)
{
    const F4 a = fluidAcceleration(sc, st, vi, doVisc, doST, nuLam, A);
    if (wallsInReach(c, sc, b, t, ri.z))
    {
        addBoundaryAndFinish(c, sc, st, b, q, vi, ri, epsH2, t, a);
    }
    else
    {
        storeAndKick(c, sc, st, vi, ri.x, t, a.x, a.y, a.z);
    }
}

// The list kernels: particles with walls within reach store g plus the fluid
// terms and join the wall queue (kBoundaryForce adds the boundary terms, in
// the same order, and kicks); the others finish here. Keeps the boundary walk
// out of the list loops' register budget.
__device__ __forceinline__ void finishForceQueued(const Ctx& c, const SceneDev& sc, const SceneState& st,
                                                  const CellBox& b, const F4& vi, const F4& ri, bool doVisc,
                                                  bool doST, float nuLam, uint32_t t, const PairAcc& A)
{
    const F4 a = fluidAcceleration(sc, st, vi, doVisc, doST, nuLam, A);
    if (wallsInReach(c, sc, b, t, ri.z))
    {
        c.acc[c.sortedVals[t]] = a;
        c.wallQueue[atomicAdd(c.wallQueueCount, 1U)] = t;
    }
    else
    {
        storeAndKick(c, sc, st, vi, ri.x, t, a.x, a.y, a.z);
    }
}

// (rho, p / rho^2, h, kWDCoef / h^4) of sorted slot t
__device__ __forceinline__ F4 ownRecord(const Ctx& c, uint32_t t)
{
    const F4 hm = c.shm[t];
    return F4 {c.fb[t].w, c.fa[t].w, hm.x, hm.w};
}

// Fluid pair sums of slot t over its neighbour list, in list order. With
// kUniform every neighbour has h0 and m0 (two gathers per pair), otherwise
// the mass and smoothing length are gathered too.
// With kSharedH (uniform scenes whose h0 lies in [2^-40, 2^39]) the quotient
// d / h0 is taken through the reciprocal of h0.
template <bool kUniform, bool kSharedH = false, bool kLam = false, bool kViscNoST = false>
__device__ __forceinline__ void forceList(const Ctx& c, const SceneDev& sc, uint32_t t, uint32_t lp, uint32_t cnt,
                                          const F4& q,
                                          const F4& vi, float rhoi, float prri, float nuNum, float epsH2, bool doVisc,
                                          bool doST, float nuLam, PairAcc& A)
{
    const float m0 = sc.m0;
    const float h0 = sc.h0;
    const float wd0 = kUniform ? dcoefOverH4(h0) : 0.0F;
    const float y0 = kSharedH ? divRecip(h0) : 0.0F;
    constexpr uint32_t B = kUniform ? kForceUniformBatch : kForceBatch;
    const uint4* row0 = reinterpret_cast<const uint4*>(c.list) + listChunk(c, lp, 0U);
    uint32_t e = 0;
    for (; e + B <= cnt; e += B)
    {
        uint32_t j[B];
        F4 fa[B];
        F4 fb[B];
        F4 hm[B];
#pragma unroll
        for (uint32_t u = 0; u < B; u += 4U)
        {
            const uint4 v = loadChunk(listRowAt(row0, e + u));
            j[u] = v.x;
            j[u + 1U] = v.y;
            j[u + 2U] = v.z;
            j[u + 3U] = v.w;
        }
#pragma unroll
        for (uint32_t u = 0; u < B; ++u)
        {
            fa[u] = recAt<F4>(c.fa, j[u]);
            fb[u] = recAt<F4>(c.fb, j[u]);
            if (!kUniform)
            {
                hm[u] = recAt<F4>(c.shm, j[u]);
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < B; ++u)
        {
            forcePair<kSharedH, kLam, kViscNoST>(fa[u], fb[u], kUniform ? m0 : hm[u].y, kUniform ? h0 : hm[u].x,
                                kUniform ? wd0 : hm[u].w, q, vi, rhoi, prri, nuNum, epsH2, doVisc, doST, nuLam, A, y0,
                                j[u] != t);
        }
    }
    for (; e < cnt; ++e)
    {
        const uint32_t jj = listRowAt(row0, e)[e % kChunk];
        if (jj != t)
        {
            if (kUniform)
            {
                forcePair<kSharedH, kLam, kViscNoST>(recAt<F4>(c.fa, jj), recAt<F4>(c.fb, jj), m0, h0, wd0, q, vi, rhoi, prri, nuNum, epsH2, doVisc, doST,
                                          nuLam, A,
                                    y0);
            }
            else
            {
                const F4 h1 = recAt<F4>(c.shm, jj);
                forcePair<false, kLam, kViscNoST>(recAt<F4>(c.fa, jj), recAt<F4>(c.fb, jj), h1.y, h1.x, h1.w, q, vi, rhoi, prri, nuNum, epsH2, doVisc,
                                       doST, nuLam, A);
            }
        }
    }
}

// Tolerance mode (kSceneToleranceMath, no surface tension or Morris term):
// computePressureAccelerations and computeViscosityAccelerations over the
// list in one sum. With r = x_i - x_j, q = d / h_j and t = 1 - q/2 the
// kernel gradient along r^ is W'(d, h_j) r / d = (kWDCoef / h_j^4) (t^3 / h_j) r,
// so both pair terms are multiples of r with the factor
// m_j (kWDCoef / h_j^5) t^3 (p_i / rho_i^2 + p_j / rho_j^2 + Pi_ij),
// Pi_ij = -2 alpha h_i c (v_ij . r) / ((rho_i + rho_j)(d^2 + 0.01 h_i^2)) for
// approaching pairs (v_ij . r < 0) and 0 otherwise. The particle itself
// (r = 0) adds a zero multiple of r. The sum goes to A.apx..apz.
template <bool kUniform>
__device__ __forceinline__ void forceListFast(const Ctx& c, const SceneDev& sc, uint32_t lp, uint32_t cnt,
                                              const F4& q, const F4& vi, float rhoi, float prri, float nuNum,
                                              float epsH2, PairAcc& A)
{
    const float invh0 = kUniform ? 1.0F / sc.h0 : 0.0F;
    const float mwd0 = kUniform ? (sc.m0 * dcoefOverH4(sc.h0)) * invh0 : 0.0F;
    const float negNu = -nuNum;
    constexpr uint32_t B = kUniform ? ASPH_FORCE_TOL_BATCH : kForceBatch;
    const uint4* row0 = reinterpret_cast<const uint4*>(c.list) + listChunk(c, lp, 0U);
    float ax = 0.0F;
    float ay = 0.0F;
    float az = 0.0F;
    for (uint32_t e = 0; e < cnt; e += B)
    {
        uint32_t j[B];
        F4 fa[B];
        F4 fb[B];
        F4 hm[B];
#pragma unroll
        for (uint32_t u = 0; u < B; u += 4U)
        {
            // chunks past the list (the tail of the last batch) are not read:
            // the list holds listCap entries, a multiple of 4 but not always of B
            const uint4 v = loadChunk(listRowAt(row0, e + u < cnt ? e + u : e));
            j[u] = v.x;
            j[u + 1U] = v.y;
            j[u + 2U] = v.z;
            j[u + 3U] = v.w;
        }
#pragma unroll
        for (uint32_t u = 0; u < B; ++u)
        {
            // entries past cnt (the tail of the last batch) read the first
            // entry again; their terms are discarded below
            const uint32_t jj = e + u < cnt ? j[u] : j[0];
            fa[u] = recAt<F4>(c.fa, jj);
            fb[u] = recAt<F4>(c.fb, jj);
            if (!kUniform)
            {
                hm[u] = recAt<F4>(c.shm, jj);
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < B; ++u)
        {
            const float ox = q.x - fa[u].x;
            const float oy = q.y - fa[u].y;
            const float oz = q.z - fa[u].z;
            const float d2 = __builtin_fmaf(oz, oz, __builtin_fmaf(oy, oy, ox * ox));
            const float invh = kUniform ? invh0 : __builtin_amdgcn_rcpf(hm[u].x);
            const float qh = fastDistOverH(d2, invh);
            const float t = __builtin_fmaf(-0.5F, qh, 1.0F);
            const float t3 = (t * t) * t;
            const float vx = vi.x - fb[u].x;
            const float vy = vi.y - fb[u].y;
            const float vz = vi.z - fb[u].z;
            const float comp = __builtin_fmaf(vz, oz, __builtin_fmaf(vy, oy, vx * ox));
            const float piv = (negNu * comp) * __builtin_amdgcn_rcpf((rhoi + fb[u].w) * (d2 + epsH2));
            const float coef = (prri + fa[u].w) + (comp < 0.0F ? piv : 0.0F);
            const float mwd = kUniform ? mwd0 : (hm[u].y * hm[u].w) * invh;
            const float sraw = (mwd * t3) * coef;
            const float sp = e + u < cnt ? sraw : 0.0F;
            ax = __builtin_fmaf(-sp, ox, ax);
            ay = __builtin_fmaf(-sp, oy, ay);
            az = __builtin_fmaf(-sp, oz, az);
        }
    }
    A.apx = ax;
    A.apy = ay;
    A.apz = az;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void kForce(Ctx c)
{
    const uint32_t n = *c.nTotal;
    const uint32_t K = c.listCap;
    // query positions pq, in the list builder's order (slot t's list at pq)
    for (uint32_t pq = xcdBlock(n) * blockDim.x + threadIdx.x; pq < n; pq += gridDim.x * blockDim.x)
    {
        const uint32_t t = querySlot(c, pq);
        const uint32_t cnt = c.ncount[t];
        if (cnt > K)
        {
            continue;  // kForceOverflow
        }
        const uint32_t s = c.sscene[t];
        const SceneDev& sc = c.scenes[s];
        if (toleranceForce(sc))
        {
            continue;  // kForceTol
        }
        const SceneState& st = c.states[s];
        const F4 q = c.s0[t];
        const F4 vi = c.s1[t];
        const F4 ri = ownRecord(c, t);
        const float hi = ri.z;
        const float rhoi = ri.x;
        const float prri = ri.y;
        const bool doVisc = sc.alpha != 0.0F;
        const bool doST = sc.sigma != 0.0F;
        const float nuLam = sc.nuLam;
        const float nuNum = ((2.0F * sc.alpha) * hi) * sc.c0;
        const float epsH2 = (0.01F * hi) * hi;
        const CellBox b = cellBox(sc, q.x, q.y, q.z, 2.0F * hi);
        PairAcc A {};
        if ((sc.flags & kSceneUniformHM) != 0U)
        {
            // (the flag is set only for h0 in [2^-40, 2^39])
            if (nuLam != 0.0F)
            {
                forceList<true, true, true>(c, sc, t, pq, cnt, q, vi, rhoi, prri, nuNum, epsH2, doVisc, doST, nuLam, A);
            }
            else if (doVisc && !doST)
            {
                forceList<true, true, false, true>(c, sc, t, pq, cnt, q, vi, rhoi, prri, nuNum, epsH2, doVisc, doST,
                                                   nuLam, A);
            }
            else
            {
                forceList<true, true>(c, sc, t, pq, cnt, q, vi, rhoi, prri, nuNum, epsH2, doVisc, doST, nuLam, A);
            }
        }
        else
        {
            if (nuLam != 0.0F)
            {
                forceList<false, false, true>(c, sc, t, pq, cnt, q, vi, rhoi, prri, nuNum, epsH2, doVisc, doST, nuLam, A);
            }
            else if (doVisc && !doST)
            {
                forceList<false, false, false, true>(c, sc, t, pq, cnt, q, vi, rhoi, prri, nuNum, epsH2, doVisc, doST,
                                                     nuLam, A);
            }
            else
            {
                forceList<false>(c, sc, t, pq, cnt, q, vi, rhoi, prri, nuNum, epsH2, doVisc, doST, nuLam, A);
            }
        }
        finishForceQueued(c, sc, st, b, vi, ri, doVisc, doST, nuLam, t, A);
    }
}

// The particles of tolerance-mode scenes (toleranceForce): forceListFast,
// then the boundary terms and the second half kick as in kForce. A kernel of
// its own so that its register budget (and occupancy) is not the exact
// loops'.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ASPH_FORCE_TOL_WAVES))) void kForceTol(Ctx c)
{
    const uint32_t n = *c.nTotal;
    const uint32_t K = c.listCap;
    for (uint32_t pq = xcdBlock(n) * blockDim.x + threadIdx.x; pq < n; pq += gridDim.x * blockDim.x)
    {
        const uint32_t t = querySlot(c, pq);
        const uint32_t cnt = c.ncount[t];
        if (cnt > K)
        {
            continue;  // kForceOverflow
        }
        const uint32_t s = c.sscene[t];
        const SceneDev& sc = c.scenes[s];
        if (!toleranceForce(sc))
        {
            continue;  // kForce
        }
        const SceneState& st = c.states[s];
        const F4 q = c.s0[t];
        const F4 vi = c.s1[t];
        const F4 ri = ownRecord(c, t);
        const float hi = ri.z;
        const float nuNum = ((2.0F * sc.alpha) * hi) * sc.c0;
        const float epsH2 = (0.01F * hi) * hi;
        const CellBox b = cellBox(sc, q.x, q.y, q.z, 2.0F * hi);
        PairAcc A {};
        if ((sc.flags & kSceneUniformHM) != 0U)
        {
            forceListFast<true>(c, sc, pq, cnt, q, vi, ri.x, ri.y, nuNum, epsH2, A);
        }
        else
        {
            forceListFast<false>(c, sc, pq, cnt, q, vi, ri.x, ri.y, nuNum, epsH2, A);
        }
        // the viscosity is inside A.ap*: no separate term
        finishForceQueued(c, sc, st, b, vi, ri, false, false, 0.0F, t, A);
    }
}

// Boundary pressure and friction of the wall queue's particles, added to the
// stored g + fluid terms, then the store and kick.
__global__ __launch_bounds__(256) void kBoundaryForce(Ctx c)
{
    const uint32_t m = *c.wallQueueCount;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x)
    {
        const uint32_t t = c.wallQueue[k];
        const uint32_t s = c.sscene[t];
        const SceneDev& sc = c.scenes[s];
        const SceneState& st = c.states[s];
        const F4 q = c.s0[t];
        const F4 vi = c.s1[t];
        const F4 ri = ownRecord(c, t);
        const float hi = ri.z;
        const float epsH2 = (0.01F * hi) * hi;
        const CellBox b = cellBox(sc, q.x, q.y, q.z, 2.0F * hi);
        addBoundaryAndFinish(c, sc, st, b, q, vi, ri, epsH2, t, c.acc[c.sortedVals[t]]);
    }
}

// Forces of the particles whose neighbours overflow their list: the walk
// again, one wave per particle (waveWalkFold), in the same order. The twelve
// accumulators are folded by twelve lanes at once: each accepting lane writes
// its pair terms to a row of the wave's LDS table (the other lanes write
// zeros, which leave a sum unchanged), and lane k adds column k row by row.
__global__ __launch_bounds__(256) void kForceOverflow(Ctx c)
{
    constexpr uint32_t kTerms = sizeof(PairAcc) / sizeof(float);
    static_assert(kTerms == 12U, "PairAcc holds twelve accumulators");
    __shared__ float4 terms[4U * 64U * kTerms / 4U];  // (float4: 16-byte rows of four terms)
    const uint32_t m = *c.overflowCount;
    const uint32_t lane = threadIdx.x & 63U;
    float* table = reinterpret_cast<float*>(terms) + (threadIdx.x / 64U) * 64U * kTerms;
    const float* column = table + (lane < kTerms ? lane : kTerms - 1U);
    const uint32_t waves = gridDim.x * (blockDim.x / 64U);
    for (uint32_t k = blockIdx.x * (blockDim.x / 64U) + threadIdx.x / 64U; k < m; k += waves)
    {
        const uint32_t t = c.overflowList[k];
        const uint32_t s = c.sscene[t];
        const SceneDev& sc = c.scenes[s];
        const SceneState& st = c.states[s];
        const F4 q = c.s0[t];
        const F4 vi = c.s1[t];
        const F4 ri = ownRecord(c, t);
        const float hi = ri.z;
        const bool doVisc = sc.alpha != 0.0F;
        const bool doST = sc.sigma != 0.0F;
        const float nuLam = sc.nuLam;
        const float nuNum = ((2.0F * sc.alpha) * hi) * sc.c0;
        const float epsH2 = (0.01F * hi) * hi;
        const CellBox b = cellBox(sc, q.x, q.y, q.z, 2.0F * hi);
        // each lane's pair terms: forcePair on zero accumulators gives -t for
        // the terms it subtracts and t for those it adds, and acc + (-t) is
        // acc - t exactly (the accumulators start at +0 and never become -0)
        PairAcc Z {};
        float sum = 0.0F;  // accumulator `lane` of PairAcc (lanes 0-11)
        waveWalkFold(
            c, sc, st.lostEnd, b, q, lane,
            [&](uint32_t j, float) {
                Z = PairAcc {};
                if (j != t)
                {
                    const F4 hm = c.shm[j];
                    if (nuLam != 0.0F)
                    {
                        forcePair<false, true>(c.fa[j], c.fb[j], hm.y, hm.x, hm.w, q, vi, ri.x, ri.y, nuNum, epsH2,
                                               doVisc, doST, nuLam, Z);
                    }
                    else
                    {
                        forcePair(c.fa[j], c.fb[j], hm.y, hm.x, hm.w, q, vi, ri.x, ri.y, nuNum, epsH2, doVisc, doST,
                                  nuLam, Z);
                    }
                }
            },
            [&](uint64_t mask) {
                const bool mine = ((mask >> lane) & 1ULL) != 0ULL;
                const float* z = reinterpret_cast<const float*>(&Z);
                float* row = table + lane * kTerms;
                waveSync();  // the previous step's readers are done
#pragma unroll
                for (uint32_t k = 0; k < kTerms; k += 4U)
                {
                    *reinterpret_cast<float4*>(row + k) = mine ? float4 {z[k], z[k + 1U], z[k + 2U], z[k + 3U]}
                                                               : float4 {0.0F, 0.0F, 0.0F, 0.0F};
                }
                waveSync();
                const uint32_t first = static_cast<uint32_t>(__builtin_ctzll(mask));
                const uint32_t last = 63U - static_cast<uint32_t>(__builtin_clzll(mask));
#pragma unroll 8
                for (uint32_t a = first; a <= last; ++a)
                {
                    sum = sum + column[a * kTerms];
                }
            });
        PairAcc A;
        float* acc = reinterpret_cast<float*>(&A);
#pragma unroll
        for (uint32_t k = 0; k < kTerms; ++k)
        {
            acc[k] = laneValue(sum, k);
        }
        if (lane == 0U)
        {
            finishForce(c, sc, st, b, q, vi, ri, epsH2, doVisc, doST, nuLam, t, A);
        }
    }
}

// ----------------------------------------------------------- refinement --

// Walk of the fluid grid of the last sort (reference order) with canonical
// arrays read through the sorted indices, exactly as forEachFluidNeighbor
// reads particle arrays through particleArrayIndices (NeighborGrid.cuh:117-121).
template <class F>
__device__ __forceinline__ void walkCanonical(const Ctx& c, const SceneDev& sc, uint32_t lost, float px, float py,
                                              float pz, float R, F&& f)
{
    const CellBox b = cellBox(sc, px, py, pz, R);
    const int gxgy = sc.gx * sc.gy;
    for (int x = b.x0; x <= b.x1; ++x)
    {
        for (int y = b.y0; y <= b.y1; ++y)
        {
            for (int z = b.z0; z <= b.z1; ++z)
            {
                const uint32_t cell = sc.cellBase + static_cast<uint32_t>(x + y * sc.gx + z * gxgy);
                const uint32_t lo = max(c.cellOff[cell], lost);
                const uint32_t hiS = c.cellOff[cell + 1U];
                for (uint32_t slot = lo; slot < hiS; ++slot)
                {
                    f(c.sortedVals[slot]);
                }
            }
        }
    }
}

// walkCanonical's cells and order, passing the sorted slot.
template <class F>
__device__ __forceinline__ void walkSlots(const Ctx& c, const SceneDev& sc, uint32_t lost, float px, float py,
                                          float pz, float R, F&& f)
{
    const CellBox b = cellBox(sc, px, py, pz, R);
    const int gxgy = sc.gx * sc.gy;
    for (int x = b.x0; x <= b.x1; ++x)
    {
        for (int y = b.y0; y <= b.y1; ++y)
        {
            for (int z = b.z0; z <= b.z1; ++z)
            {
                const uint32_t cell = sc.cellBase + static_cast<uint32_t>(x + y * sc.gx + z * gxgy);
                const uint32_t lo = max(c.cellOff[cell], lost);
                const uint32_t hiS = c.cellOff[cell + 1U];
                for (uint32_t slot = lo; slot < hiS; ++slot)
                {
                    f(slot);
                }
            }
        }
    }
}

// VorticityCriterion.cuh:67-113 (both generators)
__device__ bool vorticityMag(const Ctx& c, const SceneDev& sc, uint32_t lost, uint32_t id, float* mag)
{
    const F4 q = c.pos[id];
    const float hh = c.h[id];
    float dxx = 0.F, dxy = 0.F, dxz = 0.F, dyx = 0.F, dyy = 0.F, dyz = 0.F, dzx = 0.F, dzy = 0.F, dzz = 0.F;
    int valid = 0;
    walkCanonical(c, sc, lost, q.x, q.y, q.z, hh, [&](uint32_t j) {
        if (j == id)
        {
            return;
        }
        const F4 pj = c.pos[j];
        const float rx = pj.x - q.x;
        const float ry = pj.y - q.y;
        const float rz = pj.z - q.z;
        const float dist = sqrtf(dot3(rx, ry, rz, rx, ry, rz));
        if (dist < hh && dist > 1e-6F)
        {
            const F4 vj = c.vel[j];
            const F4 vid = c.vel[id];
            const float vdx = vj.x - vid.x;
            const float vdy = vj.y - vid.y;
            const float vdz = vj.z - vid.z;
            const float gw = wendlandGrad(dist, hh);
            const float w = c.m[j] / c.rho[j];
            const float gx = rx / dist;
            const float gy = ry / dist;
            const float gz = rz / dist;
            dxx = dxx + ((w * vdx) * gx) * gw;
            dxy = dxy + ((w * vdx) * gy) * gw;
            dxz = dxz + ((w * vdx) * gz) * gw;
            dyx = dyx + ((w * vdy) * gx) * gw;
            dyy = dyy + ((w * vdy) * gy) * gw;
            dyz = dyz + ((w * vdy) * gz) * gw;
            dzx = dzx + ((w * vdz) * gx) * gw;
            dzy = dzy + ((w * vdz) * gy) * gw;
            dzz = dzz + ((w * vdz) * gz) * gw;
            valid++;
        }
    });
    (void)dxx;
    (void)dyy;
    (void)dzz;
    if (valid < 4)
    {
        return false;
    }
    const float ox = dzy - dyz;
    const float oy = dxz - dzx;
    const float oz = dyx - dxy;
    *mag = sqrtf(dot3(ox, oy, oz, ox, oy, oz));
    return true;
}

// checkIfParticleCanBeSplitInsideDomain (criteria/Common.cuh:6-20)
__device__ __forceinline__ bool inDomain(const SceneDev& sc, const F4& q, float hh)
{
    const float mo = sc.eps * hh;
    return !(q.x - mo < sc.dmin[0] || q.x + mo > sc.dmax[0] || q.y - mo < sc.dmin[1] || q.y + mo > sc.dmax[1] ||
             q.z - mo < sc.dmin[2] || q.z + mo > sc.dmax[2]);
}

// calculateMinSurfaceDistance (InterfaceCriterion.cuh:52-67)
__device__ __forceinline__ void minSurface(const SceneDev& sc, const F4& q, float* md)
{
    const float mx = gmin(q.x - sc.dmin[0], sc.dmax[0] - q.x);
    const float my = gmin(q.y - sc.dmin[1], sc.dmax[1] - q.y);
    const float mz = gmin(q.z - sc.dmin[2], sc.dmax[2] - q.z);
    md[0] = gmax(0.0F, mx);
    md[1] = gmax(0.0F, my);
    md[2] = gmax(0.0F, mz);
}

// Split generators (VelocityCriterion.cuh:54-78, InterfaceCriterion.cuh:69-95,
// VorticityCriterion.cuh:53-125) and the zone extension.
__device__ float splitCriterion(const Ctx& c, const SceneDev& sc, const SceneState& st, uint32_t id)
{
    const float minMass = sc.minMassRatio * sc.m0;
    if (c.m[id] < minMass)
    {
        return -1.0F;
    }
    const F4 q = c.pos[id];
    const float hh = c.h[id];
    if (sc.criterion == 1U)
    {
        float md[3];
        minSurface(sc, q, md);
        const float mo = sc.eps * hh;
        if (gmin(gmin(md[0], md[1]), md[2]) < mo)
        {
            return -1.0F;
        }
        const float tx = sc.ifSplit * (sc.dmax[0] - sc.dmin[0]);
        const float ty = sc.ifSplit * (sc.dmax[1] - sc.dmin[1]);
        const float tz = sc.ifSplit * (sc.dmax[2] - sc.dmin[2]);
        return 1.0F - gmin(gmin(md[0] / tx, md[1] / ty), md[2] / tz);
    }
    if (!inDomain(sc, q, hh))
    {
        return -1.0F;
    }
    if (sc.criterion == 2U)
    {
        float mag = 0.F;
        if (!vorticityMag(c, sc, st.lostEnd, id, &mag))
        {
            return -1.0F;
        }
        return mag > sc.vortSplit ? mag : -1.0F;
    }
    if (sc.criterion == 3U)
    {
        const float rx = q.x - st.zoneNow[0];
        const float ry = q.y - st.zoneNow[1];
        const float rz = q.z - st.zoneNow[2];
        const float d = sqrtf(dot3(rx, ry, rz, rx, ry, rz));
        return d < sc.zoneR ? 2.0F - d / sc.zoneR : -1.0F;
    }
    const F4 v = c.vel[id];
    const float mag = sqrtf(dot3(v.x, v.y, v.z, v.x, v.y, v.z));
    if (mag < sc.velSplit)
    {
        return -1.0F;
    }
    return mag;
}

// Merge generators (VelocityCriterion.cuh:80-101, InterfaceCriterion.cuh:97-115,
// VorticityCriterion.cuh:127-195) and the zone extension.
__device__ float mergeCriterion(const Ctx& c, const SceneDev& sc, const SceneState& st, uint32_t id)
{
    const float maxMass = sc.maxMassRatio * sc.m0;
    if (c.m[id] > maxMass)
    {
        return -1.0F;
    }
    const F4 q = c.pos[id];
    if (sc.criterion == 1U)
    {
        float md[3];
        minSurface(sc, q, md);
        const float tx = sc.ifMerge * (sc.dmax[0] - sc.dmin[0]);
        const float ty = sc.ifMerge * (sc.dmax[1] - sc.dmin[1]);
        const float tz = sc.ifMerge * (sc.dmax[2] - sc.dmin[2]);
        return gmin(gmin(md[0] / tx, md[1] / ty), md[2] / tz) - 1.0F;
    }
    if (sc.criterion == 2U)
    {
        float mag = 0.F;
        if (!vorticityMag(c, sc, st.lostEnd, id, &mag))
        {
            return -1.0F;
        }
        return mag < sc.vortMerge ? mag / sc.vortMerge : -1.0F;
    }
    if (sc.criterion == 3U)
    {
        const float rx = q.x - st.zoneNow[0];
        const float ry = q.y - st.zoneNow[1];
        const float rz = q.z - st.zoneNow[2];
        const float d = sqrtf(dot3(rx, ry, rz, rx, ry, rz));
        return d > sc.zoneR ? sc.zoneR / d : 2.0F;
    }
    const F4 v = c.vel[id];
    const float mag = sqrtf(dot3(v.x, v.y, v.z, v.x, v.y, v.z));
    if (mag < sc.velMerge)
    {
        return mag / sc.velMerge;
    }
    return -1.0F;
}

// *ctr += pred for the active lanes, where ctr is a per-scene counter of
// scene s: one atomic per wave when every active lane belongs to the same
// scene (the usual case), one per lane otherwise. Counting only, so the
// result does not depend on the order of the additions.
__device__ __forceinline__ void countAdd(uint32_t* ctr, uint32_t s, bool pred)
{
    const uint32_t s0 = __builtin_amdgcn_readfirstlane(s);
    if (__ballot((s != s0) ? 1 : 0) == 0ULL)
    {
        const uint64_t m = __ballot(pred ? 1 : 0);
        const uint64_t active = __ballot(1);
        if (m != 0ULL && __lane_id() == static_cast<uint32_t>(__ffsll(static_cast<long long>(active)) - 1))
        {
            atomicAdd(ctr, static_cast<uint32_t>(__popcll(m)));
        }
    }
    else if (pred)
    {
        atomicAdd(ctr, 1U);
    }
}


// getCriterionValues (AdaptiveAlgorithm.cuh:40-54) for the merge criteria;
// keys (scene, value ascending) for the stable sort of findTopParticlesToMerge.
__global__ __launch_bounds__(256) void kMergeCriteria(Ctx c)
{
    const uint32_t n = *c.nTotal;
    for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < n; d += gridDim.x * blockDim.x)
    {
        const uint32_t s = segmentOf(c.denseStarts, c.sceneSlots, d);
        const SceneDev& sc = c.scenes[s];
        SceneState& st = c.states[s];
        if (!mergeParticipant(sc, st))
        {
            continue;
        }
        const uint32_t l = d - st.denseOff;
        c.cand[d] = 0U;  // reset value of mergeCandidates (AdaptiveSphSimulation.cu:127)
        c.removal[d] = 0U;
        c.keep[d] = 0U;
        const float v = mergeCriterion(c, sc, st, sc.pOff + l);
        const uint32_t k = st.segStart + l;  // the merge layout of kPrep
        c.ckeysA[k] = (static_cast<uint64_t>(s) << 32U) | static_cast<uint64_t>(orderedBits(v));
        c.cvalsA[k] = l;
        countAdd(&st.positives, s, v > 0.0F);
    }
}

// count_if + min(count, round(N * maxBatchRatio)) (AdaptiveAlgorithm.cuh:109-128)
__global__ void kMergeSelect(Ctx c)
{
    if (threadIdx.x != 0)
    {
        return;
    }
    uint32_t run = 0;
    for (uint32_t s = 0; s < c.sceneSlots; ++s)
    {
        const SceneDev& sc = c.scenes[s];
        SceneState& st = c.states[s];
        if (!mergeParticipant(sc, st))
        {
            continue;
        }
        st.segStart = run;
        run += st.countBefore;
        const float limit = roundf(static_cast<float>(st.countBefore) * sc.batchRatio);
        st.mergeK = min(st.positives, static_cast<uint32_t>(limit));
    }
}

__global__ void kCopyMergeList(Ctx c)
{
    const SceneState& st = c.states[0];
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < st.mergeK; e += gridDim.x * blockDim.x)
    {
        c.mergeListDbg[e] = c.csortedVals[st.segStart + e];
    }
}

// identifyMergeCandidates (AdaptiveAlgorithm.cu:159-206): nearest neighbour
// within min(h_i, h_j) (search radius h_i) lighter than maxMassRatio * m0;
// the first strictly closer hit in walk order wins.
__global__ __launch_bounds__(256) void kMergeCandidates(Ctx c)
{
    const uint32_t n = *c.critCount;  // the sorted merge keys
    const uint32_t lane = threadIdx.x & 63U;
    // waves of 64 consecutive keys (so that a wave can walk for one lane)
    for (uint32_t t0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63U); t0 < n; t0 += gridDim.x * blockDim.x)
    {
        const uint32_t t = t0 + lane;
        const uint64_t key = t < n ? c.csortedKeys[t] : ~0ULL;
        const uint32_t s = key == ~0ULL ? 0U : static_cast<uint32_t>(key >> 32U);
        const SceneDev& sc = c.scenes[s];
        const SceneState& st = c.states[s];
        const uint32_t e = t - st.segStart;
        const bool active = key != ~0ULL && e < st.mergeK;
        const uint32_t id = active ? c.csortedVals[t] : 0U;
        const uint32_t i = sc.pOff + id;
        const F4 q = active ? c.pos[i] : F4 {};
        const float hh = active ? c.h[i] : 0.0F;
        const float hh2 = hh * hh;
        const float massLimit = sc.maxMassRatio * sc.m0;
        float closest = FLT_MAX;
        uint32_t bestSlot = UINT32_MAX;
        // The search reads the sorted shadows, which hold the candidates'
        // current positions, smoothing lengths and masses (nothing changes
        // them between the gather and the merge), in the reference's order.
        // d2 < hmin^2 implies d2 < h_i^2, so the shadow's h and m are read only
        // past that test, and the candidate's identity only on acceptance.
        const uint32_t ti = active ? c.slotOf[i] : 0U;
        // d2 of slot when it is a closer admissible candidate, else -1
        auto visit = [&](uint32_t slot) {
            const F4 o = c.s0[slot];
            const float ox = o.x - q.x;
            const float oy = o.y - q.y;
            const float oz = o.z - q.z;
            const float d2 = dot3w(ox, oy, oz, ox, oy, oz);
            if (d2 < hh2)
            {
                const F4 hm = c.shm[slot];
                const float hmin = gmin(hh, hm.x);
                if (d2 < hmin * hmin && d2 < closest && hm.y < massLimit && slot != ti)
                {
                    return d2;
                }
            }
            return -1.0F;
        };
        const uint32_t cnt = active ? c.ncount[ti] : 0U;
        if (active && cnt <= c.listCap)
        {
            const uint32_t lpi = queryOrdered(c, c.sscene[ti]) ? c.qpos[ti] : ti;  // the list's query position
            // Every hit of the walk over the cells within h_i is in the
            // particle's neighbour list of this step (d2 < h_j^2 < (2 h_j)^2,
            // cells within h_i inside those within 2 h_i), and the list holds
            // the walk order (cells x, y, z, ascending slot) of a larger cell
            // box: the list, restricted to the cells within h_i, is the walk.
            const CellBox b = cellBox(sc, q.x, q.y, q.z, hh);
            for (uint32_t k = 0; k < cnt; ++k)
            {
                const uint32_t slot = *listAt(c, lpi, k);
                const float d2 = visit(slot);
                if (d2 >= 0.0F)
                {
                    const int local = static_cast<int>(c.sortedKeys[slot] - sc.cellBase);
                    const int x = local % sc.gx;
                    const int y = (local / sc.gx) % sc.gy;
                    const int z = local / (sc.gx * sc.gy);
                    if (x >= b.x0 && x <= b.x1 && y >= b.y0 && y <= b.y1 && z >= b.z0 && z <= b.z1)
                    {
                        closest = d2;
                        bestSlot = slot;
                    }
                }
            }
        }
        // Lists that overflowed: the walk over the cells within h_i
        // (walkSlots' cells and order), by the whole wave for one lane at a
        // time, 64 slots per step; the admissible slots of a step are then
        // taken in ascending order, a closer one replacing the best.
        uint64_t walkers = __ballot((active && cnt > c.listCap) ? 1 : 0);
        while (walkers != 0ULL)
        {
            const int l = static_cast<int>(__builtin_ctzll(walkers));
            walkers &= walkers - 1ULL;
            const uint32_t sl = __builtin_amdgcn_readlane(s, l);
            const SceneDev& scl = c.scenes[sl];
            const F4 ql {laneValue(q.x, l), laneValue(q.y, l), laneValue(q.z, l), 0.0F};
            const float hl = laneValue(hh, l);
            const float hl2 = hl * hl;
            const uint32_t tl = __builtin_amdgcn_readlane(ti, l);
            const float limitl = scl.maxMassRatio * scl.m0;
            const uint32_t lost = c.states[sl].lostEnd;
            const CellBox b = cellBox(scl, ql.x, ql.y, ql.z, hl);
            ConstU32* off = (ConstU32*)(c.cellOff);
            const int gxgy = scl.gx * scl.gy;
            float best = FLT_MAX;
            uint32_t bestAt = UINT32_MAX;
            for (int x = b.x0; x <= b.x1; ++x)
            {
                for (int y = b.y0; y <= b.y1; ++y)
                {
                    for (int z = b.z0; z <= b.z1; ++z)
                    {
                        const uint32_t cell = scl.cellBase + static_cast<uint32_t>(x + y * scl.gx + z * gxgy);
                        const uint32_t lo = __builtin_amdgcn_readfirstlane(max(off[cell], lost));
                        const uint32_t hi = __builtin_amdgcn_readfirstlane(off[cell + 1U]);
                        for (uint32_t j0 = lo; j0 < hi; j0 += 64U)
                        {
                            const uint32_t j = j0 + lane;
                            const bool in = j < hi;
                            const F4 o = c.s0[in ? j : hi - 1U];
                            const float ox = o.x - ql.x;
                            const float oy = o.y - ql.y;
                            const float oz = o.z - ql.z;
                            const float d2 = dot3w(ox, oy, oz, ox, oy, oz);
                            bool adm = in && d2 < hl2;
                            if (adm)
                            {
                                const F4 hm = c.shm[j];
                                const float hmin = gmin(hl, hm.x);
                                adm = d2 < hmin * hmin && hm.y < limitl && j != tl;
                            }
                            uint64_t m = __ballot(adm ? 1 : 0);
                            while (m != 0ULL)
                            {
                                const uint32_t a = static_cast<uint32_t>(__builtin_ctzll(m));
                                m &= m - 1ULL;
                                const float da = laneValue(d2, a);
                                if (da < best)
                                {
                                    best = da;
                                    bestAt = j0 + a;
                                }
                            }
                        }
                    }
                }
            }
            if (lane == static_cast<uint32_t>(l))
            {
                closest = best;
                bestSlot = bestAt;
            }
        }
        if (active)
        {
            c.cand[st.denseOff + id] = bestSlot == UINT32_MAX ? kNoCandidate : c.sortedVals[bestSlot] - sc.pOff;
        }
    }
}

// resolveMergePairs (AdaptiveAlgorithm.cu:208-234): mutual candidates (i < c).
__global__ __launch_bounds__(256) void kResolvePairs(Ctx c)
{
    const uint32_t n = *c.nTotal;
    for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < n; d += gridDim.x * blockDim.x)
    {
        const uint32_t s = segmentOf(c.denseStarts, c.sceneSlots, d);
        const SceneDev& sc = c.scenes[s];
        SceneState& st = c.states[s];
        if (!mergeParticipant(sc, st) || st.mergeK == 0U)
        {
            continue;
        }
        const uint32_t l = d - st.denseOff;
        const uint32_t cc = c.cand[d];
        if (cc == kNoCandidate)
        {
            continue;
        }
        const bool pair = c.cand[st.denseOff + cc] == l && l < cc;
        if (pair)
        {
            c.keep[d] = 1U;
            c.removal[st.denseOff + cc] = 1U;
        }
        countAdd(&st.mergePairs, s, pair);
    }
}

// performMerges + calculateMergedSmoothingLength (AdaptiveAlgorithm.cu:50-67,236-275)
__global__ __launch_bounds__(256) void kMergeApply(Ctx c)
{
    const uint32_t n = *c.nTotal;
    for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < n; d += gridDim.x * blockDim.x)
    {
        if (c.keep[d] == 0U)
        {
            continue;
        }
        const uint32_t s = segmentOf(c.denseStarts, c.sceneSlots, d);
        const SceneDev& sc = c.scenes[s];
        const SceneState& st = c.states[s];
        if (!mergeParticipant(sc, st))
        {
            continue;
        }
        const uint32_t a = sc.pOff + (d - st.denseOff);
        const uint32_t b = sc.pOff + c.cand[d];
        const F4 pa = c.pos[a];
        const F4 pb = c.pos[b];
        const float ma = c.m[a];
        const float mb = c.m[b];
        const float ha = c.h[a];
        const float hb = c.h[b];
        const F4 va = c.vel[a];
        const F4 vb = c.vel[b];
        const F4 aa = c.acc[a];
        const F4 ab = c.acc[b];
        const float M = ma + mb;
        const F4 np {(ma * pa.x + mb * pb.x) / M, (ma * pa.y + mb * pb.y) / M, (ma * pa.z + mb * pb.z) / M,
                     (ma * pa.w + mb * pb.w) / M};
        const F4 nv {(ma * va.x + mb * vb.x) / M, (ma * va.y + mb * vb.y) / M, (ma * va.z + mb * vb.z) / M,
                     (ma * va.w + mb * vb.w) / M};
        const F4 na {(ma * aa.x + mb * ab.x) / M, (ma * aa.y + mb * ab.y) / M, (ma * aa.z + mb * ab.z) / M,
                     (ma * aa.w + mb * ab.w) / M};
        const float rax = pa.x - np.x, ray = pa.y - np.y, raz = pa.z - np.z;
        const float rbx = pb.x - np.x, rby = pb.y - np.y, rbz = pb.z - np.z;
        const float da = sqrtf(dot3(rax, ray, raz, rax, ray, raz));
        const float db = sqrtf(dot3(rbx, rby, rbz, rbx, rby, rbz));
        const float denom = (ma * wendland(da, ha)) + (mb * wendland(db, hb));
        c.pos[a] = np;
        c.vel[a] = nv;
        c.acc[a] = na;
        c.m[a] = M;
        c.rad[a] = sc.r0 * detCbrt(M / sc.m0);
        const float coef = (sc.flags & kSceneKernelMergedH) != 0U ? kWCoef : kMergeCoef;
        c.h[a] = detCbrt((coef * M) / denom);
        c.rho[a] = 0.0F;
    }
}

__device__ __forceinline__ bool compacts(const SceneDev& sc, const SceneState& st)
{
    return mergeParticipant(sc, st) && st.mergeK > 0U && st.mergePairs > 0U;
}

// compactParticles (AdaptiveAlgorithm.cu:130-157), stable and out of place.
__global__ __launch_bounds__(256) void kCompact(Ctx c)
{
    const uint32_t n = *c.nTotal;
    for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < n; d += gridDim.x * blockDim.x)
    {
        const uint32_t s = segmentOf(c.denseStarts, c.sceneSlots, d);
        const SceneDev& sc = c.scenes[s];
        const SceneState& st = c.states[s];
        if (!compacts(sc, st) || c.removal[d] != 0U)
        {
            continue;
        }
        const uint32_t l = d - st.denseOff;
        const uint32_t nl = l - (c.rprefix[d] - c.rprefix[st.denseOff]);
        const uint32_t src = sc.pOff + l;
        const uint32_t dst = sc.pOff + nl;
        c.xpos[dst] = c.pos[src];
        c.xvel[dst] = c.vel[src];
        c.xacc[dst] = c.acc[src];
        c.xm[dst] = c.m[src];
        c.xrad[dst] = c.rad[src];
        c.xh[dst] = c.h[src];
        c.xrho[dst] = c.rho[src];
    }
}

// updateParticleCount (AdaptiveAlgorithm.cu:119-128)
__global__ void kAfterMerge(Ctx c)
{
    if (threadIdx.x != 0)
    {
        return;
    }
    uint32_t run = 0;
    for (uint32_t s = 0; s < c.sceneSlots; ++s)
    {
        const SceneDev& sc = c.scenes[s];
        SceneState& st = c.states[s];
        if ((sc.flags & kSceneActive) == 0U || st.refineNow == 0U)
        {
            continue;
        }
        uint32_t removed = 0;
        if (compacts(sc, st))
        {
            const uint32_t last = st.denseOff + st.countBefore - 1U;
            removed = c.rprefix[last] + c.removal[last] - c.rprefix[st.denseOff];
        }
        st.removed = removed;
        st.countAfterMerge = st.countBefore - removed;
        st.count = st.countAfterMerge;
        st.positives = 0;
        if (removed > 0U)
        {
            // split layout: the splitting scenes' keys back to back
            st.segStart = run;
            run += st.countAfterMerge;
        }
    }
    *c.critCount = run;
}

__global__ __launch_bounds__(256) void kCopyBack(Ctx c)
{
    const uint32_t n = *c.nTotal;
    for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < n; d += gridDim.x * blockDim.x)
    {
        const uint32_t s = segmentOf(c.denseStarts, c.sceneSlots, d);
        const SceneDev& sc = c.scenes[s];
        const SceneState& st = c.states[s];
        if (st.refineNow == 0U || st.removed == 0U)
        {
            continue;
        }
        const uint32_t l = d - st.denseOff;
        if (l >= st.countAfterMerge)
        {
            continue;
        }
        const uint32_t k = sc.pOff + l;
        c.pos[k] = c.xpos[k];
        c.vel[k] = c.xvel[k];
        c.acc[k] = c.xacc[k];
        c.m[k] = c.xm[k];
        c.rad[k] = c.xrad[k];
        c.h[k] = c.xh[k];
        c.rho[k] = c.xrho[k];
    }
}

// Split criteria over the compacted particles; keys (scene, value descending).
__global__ __launch_bounds__(256) void kSplitCriteria(Ctx c)
{
    const uint32_t n = *c.nTotal;
    for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < n; d += gridDim.x * blockDim.x)
    {
        const uint32_t s = segmentOf(c.denseStarts, c.sceneSlots, d);
        const SceneDev& sc = c.scenes[s];
        SceneState& st = c.states[s];
        const uint32_t l = d - st.denseOff;
        if ((sc.flags & kSceneActive) == 0U || st.refineNow == 0U || st.removed == 0U || l >= st.countAfterMerge)
        {
            continue;
        }
        const float v = splitCriterion(c, sc, st, sc.pOff + l);
        const uint32_t k = st.segStart + l;  // the split layout of kAfterMerge
        c.ckeysA[k] = (static_cast<uint64_t>(s) << 32U) | static_cast<uint64_t>(~orderedBits(v));
        c.cvalsA[k] = l;
        countAdd(&st.positives, s, v > 0.0F);
    }
}

// findTopParticlesToSplit (AdaptiveAlgorithm.cuh:56-94) and the split count of
// identifyAndSplitParticles (AdaptiveSphSimulation.cu:173-190): the kernel
// runs ceil(toSplit / tpb) * tpb threads gated by the device count k; each
// split needs 12 free slots below maxParticleCount.
__global__ void kSplitSelect(Ctx c)
{
    if (threadIdx.x != 0)
    {
        return;
    }
    uint32_t run = 0;
    for (uint32_t s = 0; s < c.sceneSlots; ++s)
    {
        const SceneDev& sc = c.scenes[s];
        SceneState& st = c.states[s];
        if ((sc.flags & kSceneActive) == 0U || st.refineNow == 0U)
        {
            continue;
        }
        if (st.removed == 0U)
        {
            st.splitK = 0;
            st.splitDone = 0;
            st.countAfter = st.count;
            continue;
        }
        st.segStart = run;
        const uint32_t np = st.countAfterMerge;
        run += np;
        const float limit = roundf(static_cast<float>(np) * sc.batchRatio);
        const uint32_t k = min(st.positives, static_cast<uint32_t>(limit));
        st.splitK = k;
        const uint32_t toSplit = min(k, st.removed / 12U);
        uint32_t done = 0;
        if (toSplit > 0U)
        {
            const uint32_t launched = ((toSplit + sc.tpb - 1U) / sc.tpb) * sc.tpb;
            const uint32_t eff = min(k, launched);
            const uint32_t room = sc.maxCount >= np ? (sc.maxCount - np) / 12U : 0U;
            done = min(eff, room);
        }
        st.splitDone = done;
        st.count = np + 12U * done;
        st.countAfter = st.count;
    }
}

// splitParticles (AdaptiveAlgorithm.cu:69-117); daughters of the r-th parent
// take slots countAfterMerge + 12 r (the reference takes them by atomicAdd).
__global__ __launch_bounds__(256) void kSplitApply(Ctx c)
{
    constexpr float phi = 1.618033988749894848204586834365638F;
    constexpr float inv = 0.5257311121F;
    constexpr float pin = phi * inv;
    const float verts[12][3] = {{0.F, pin, inv},  {0.F, pin, -inv},  {0.F, -pin, -inv}, {0.F, -pin, inv},
                                {inv, 0.F, pin},  {-inv, 0.F, pin},  {inv, 0.F, -pin},  {-inv, 0.F, -pin},
                                {pin, inv, 0.F},  {pin, -inv, 0.F},  {-pin, inv, 0.F},  {-pin, -inv, 0.F}};
    const uint32_t n = *c.critCount;  // the sorted split keys
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    {
        const uint64_t key = c.csortedKeys[t];
        if (key == ~0ULL)
        {
            continue;
        }
        const uint32_t s = static_cast<uint32_t>(key >> 32U);
        const SceneDev& sc = c.scenes[s];
        const SceneState& st = c.states[s];
        const uint32_t e = t - st.segStart;
        if (e >= st.splitDone)
        {
            continue;
        }
        const uint32_t parent = sc.pOff + c.csortedVals[t];
        const uint32_t base = sc.pOff + st.countAfterMerge + 12U * e;
        const F4 op = c.pos[parent];
        const F4 ov = c.vel[parent];
        const F4 oa = c.acc[parent];
        const float om = c.m[parent];
        const float orad = c.rad[parent];
        const float oh = c.h[parent];
        const float dm = om * sc.vertexRatio;
        const float cm = om * sc.centerRatio;
        const float newRadius = orad * detCbrt(dm / om);
        const float newH = sc.alphaSplit * oh;
        c.m[parent] = cm;
        c.rad[parent] = orad * detCbrt(cm / om);
        c.h[parent] = newH;
        for (int v = 0; v < 12; ++v)
        {
            const uint32_t ni = base + static_cast<uint32_t>(v);
            const float ox = (verts[v][0] * sc.eps) * oh;
            const float oy = (verts[v][1] * sc.eps) * oh;
            const float oz = (verts[v][2] * sc.eps) * oh;
            c.pos[ni] = F4 {op.x + ox, op.y + oy, op.z + oz, op.w + 0.0F};
            c.vel[ni] = ov;
            c.acc[ni] = oa;
            c.m[ni] = dm;
            c.rad[ni] = newRadius;
            c.h[ni] = newH;
            c.rho[ni] = 0.0F;
        }
    }
}

// Second half kick of the scenes that refined this step (AdaptiveSphSimulation.cu:81).
__global__ __launch_bounds__(256) void kKickRefined(Ctx c)
{
    const uint32_t s = blockIdx.y;
    const SceneDev& sc = c.scenes[s];
    const SceneState& st = c.states[s];
    if ((sc.flags & kSceneActive) == 0U || st.refineNow == 0U)
    {
        return;
    }
    const uint32_t cnt = st.count;
    const float hdt = sc.halfDt;
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < cnt; l += gridDim.x * blockDim.x)
    {
        const uint32_t i = sc.pOff + l;
        F4 v = c.vel[i];
        const F4 a = c.acc[i];
        v.x = v.x + a.x * hdt;
        v.y = v.y + a.y * hdt;
        v.z = v.z + a.z * hdt;
        v.w = v.w + a.w * hdt;
        const float len = sqrtf(dot4(v.x, v.y, v.z, v.w));
        if (len > sc.vmax)
        {
            const float k = sc.vmax / len;
            v.x = v.x * k;
            v.y = v.y * k;
            v.z = v.z * k;
            v.w = v.w * k;
        }
        c.vel[i] = v;
    }
}

// ---------------------------------------------------------- diagnostics --
// sumAllNeighbors (Algorithm.cu:234-268) with per-particle counts and set
// fingerprints, over the grid of the last update and the current arrays.
__global__ __launch_bounds__(256) void kNeighborSets(Ctx c, uint32_t scene, uint32_t count, uint32_t* counts,
                                                     uint64_t* prints, uint32_t* total)
{
    const SceneDev& sc = c.scenes[scene];
    const SceneState& st = c.states[scene];
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < count; l += gridDim.x * blockDim.x)
    {
        const uint32_t i = sc.pOff + l;
        const F4 q = c.pos[i];
        uint32_t cnt = 0;
        uint64_t hsh = 0;
        walkCanonical(c, sc, st.lostEnd, q.x, q.y, q.z, 2.0F * c.h[i], [&](uint32_t j) {
            if (j == i)
            {
                return;
            }
            const F4 pj = c.pos[j];
            const float ox = pj.x - q.x;
            const float oy = pj.y - q.y;
            const float oz = pj.z - q.z;
            const float d2 = dot3w(ox, oy, oz, ox, oy, oz);
            const float hj = c.h[j];
            const float r2 = (2.0F * hj) * (2.0F * hj);
            if (d2 <= r2)
            {
                cnt++;
                hsh += mix64(static_cast<uint64_t>(j - sc.pOff));
            }
        });
        if (counts != nullptr)
        {
            counts[l] = cnt;
        }
        if (prints != nullptr)
        {
            prints[l] = hsh;
        }
        atomicAdd(total, cnt);
    }
}

// getDensityInfo partials (SphSimulation.cu:249-285): per block min, max,
// FP64 sum and the three threshold counts.
__global__ __launch_bounds__(256) void kDensityInfo(Ctx c, uint32_t scene, uint32_t count, float threshold,
                                                    double* dpart, uint32_t* upart)
{
    __shared__ float smin[256];
    __shared__ float smax[256];
    __shared__ double ssum[256];
    __shared__ uint32_t scnt[3][256];
    const SceneDev& sc = c.scenes[scene];
    float mn = FLT_MAX;
    float mx = -FLT_MAX;
    double sum = 0.0;
    uint32_t under = 0, normal = 0, over = 0;
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < count; l += gridDim.x * blockDim.x)
    {
        const float dns = c.rho[sc.pOff + l];
        const float ratio = dns / sc.rho0 - 1.0F;
        mn = fminf(mn, dns);
        mx = fmaxf(mx, dns);
        sum += static_cast<double>(dns);
        under += (ratio < -threshold) ? 1U : 0U;
        normal += (ratio >= -threshold && ratio <= threshold) ? 1U : 0U;
        over += (ratio > threshold) ? 1U : 0U;
    }
    const uint32_t tid = threadIdx.x;
    smin[tid] = mn;
    smax[tid] = mx;
    ssum[tid] = sum;
    scnt[0][tid] = under;
    scnt[1][tid] = normal;
    scnt[2][tid] = over;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1U)
    {
        if (tid < w)
        {
            smin[tid] = fminf(smin[tid], smin[tid + w]);
            smax[tid] = fmaxf(smax[tid], smax[tid + w]);
            ssum[tid] = ssum[tid] + ssum[tid + w];
            scnt[0][tid] += scnt[0][tid + w];
            scnt[1][tid] += scnt[1][tid + w];
            scnt[2][tid] += scnt[2][tid + w];
        }
        __syncthreads();
    }
    if (tid == 0)
    {
        dpart[blockIdx.x] = ssum[0];
        upart[5 * blockIdx.x + 0] = __float_as_uint(smin[0]);
        upart[5 * blockIdx.x + 1] = __float_as_uint(smax[0]);
        upart[5 * blockIdx.x + 2] = scnt[0][0];
        upart[5 * blockIdx.x + 3] = scnt[1][0];
        upart[5 * blockIdx.x + 4] = scnt[2][0];
    }
}

}  // namespace

void launchPrep(const Ctx& c, hipStream_t s) { hipLaunchKernelGGL(kPrep, dim3(1), dim3(256), 0, s, c); }

void launchKickDriftKey(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kKickDriftKey, dim3(gridFor(c.nMax, lc.blockStream)), dim3(lc.blockStream), 0, s, c);
}


void launchQueryOrder(const Ctx& c, hipStream_t s)
{
    const uint32_t blocks = (c.cExtent + 3U) / 4U;
    hipLaunchKernelGGL(kQueryOrder, dim3(blocks < kGridCap ? (blocks == 0U ? 1U : blocks) : kGridCap), dim3(256), 0, s,
                       c);
}

void launchGather(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kGather, dim3(gridFor(c.nMax, lc.blockStream)), dim3(lc.blockStream), 0, s, c);
}

// Overflow kernels: a grid that covers a queue of any length with enough
// parallelism (refined regions can overflow a large share of the particles)
// while costing little when the queue is empty.
__host__ __forceinline__ uint32_t overflowGrid(const Ctx& c)
{
    constexpr uint32_t kMaxBlocks = 2048;
    const uint32_t g = gridFor(c.nMax, 256);
    return g < kMaxBlocks ? g : kMaxBlocks;
}

void launchNeighbors(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    const uint32_t b = lc.blockNeighbor;
    if (b == 64U)
    {
        hipLaunchKernelGGL(kNeighbors<64>, dim3(gridCover(c.nMax + 64U * kHeavyCap, b)), dim3(b), 0, s, c);
    }
    else if (b == 128U)
    {
        hipLaunchKernelGGL(kNeighbors<128>, dim3(gridCover(c.nMax + 64U * kHeavyCap, b)), dim3(b), 0, s, c);
    }
    else
    {
        hipLaunchKernelGGL(kNeighbors<256>, dim3(gridCover(c.nMax + 64U * kHeavyCap, 256U)), dim3(256), 0, s, c);
    }
}

void launchDensity(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    // the list builder sums the densities of the particles whose list fits
    hipLaunchKernelGGL(kDensityOverflow, dim3(overflowGrid(c)), dim3(256), 0, s, c);
    (void)lc;
}

void launchForce(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    // one thread per slot, uncapped: particles differ widely in cost (a
    // refined region's hundreds of neighbours), and many small workgroups let
    // the dispatcher balance them where a capped grid would give each wave a
    // fixed share
    // (a persistent grid, LaunchCfg::forceBlocksPerCU, is the alternative the
    // launch-shape tuner weighs against this)
    const uint32_t b = lc.blockNeighbor;
    auto grid = [&](uint32_t bs) {
        const uint32_t g = gridCover(c.nMax, bs);
        return lc.forceGrid != 0U && lc.forceGrid < g ? lc.forceGrid : g;
    };
    if ((c.forceKinds & kForceExact) != 0U)
    {
        hipLaunchKernelGGL(kForce, dim3(grid(b)), dim3(b), 0, s, c);
    }
    if ((c.forceKinds & kForceTolerance) != 0U)
    {
        hipLaunchKernelGGL(kForceTol, dim3(grid(256U)), dim3(256), 0, s, c);
    }
    hipLaunchKernelGGL(kBoundaryForce, dim3(overflowGrid(c)), dim3(256), 0, s, c);
    hipLaunchKernelGGL(kForceOverflow, dim3(overflowGrid(c)), dim3(256), 0, s, c);
}

void launchMergeCriteria(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kMergeCriteria, dim3(gridFor(c.nMax, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchMergeSelect(const Ctx& c, hipStream_t s) { hipLaunchKernelGGL(kMergeSelect, dim3(1), dim3(64), 0, s, c); }

void launchMergeCandidates(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kMergeCandidates, dim3(gridFor(c.nMax, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchResolvePairs(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kResolvePairs, dim3(gridFor(c.nMax, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchMergeApply(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kMergeApply, dim3(gridFor(c.nMax, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchCompact(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kCompact, dim3(gridFor(c.nMax, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchAfterMerge(const Ctx& c, hipStream_t s) { hipLaunchKernelGGL(kAfterMerge, dim3(1), dim3(64), 0, s, c); }

void launchCopyBack(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kCopyBack, dim3(gridFor(c.nMax, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchSplitCriteria(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kSplitCriteria, dim3(gridFor(c.nMax, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchSplitSelect(const Ctx& c, hipStream_t s) { hipLaunchKernelGGL(kSplitSelect, dim3(1), dim3(64), 0, s, c); }

void launchSplitApply(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kSplitApply, dim3(gridFor(c.nMax, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchKickRefined(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kKickRefined, dim3(64, c.sceneSlots), dim3(256), 0, s, c);
    (void)lc;
}

void launchCopyMergeList(const Ctx& c, hipStream_t s)
{
    hipLaunchKernelGGL(kCopyMergeList, dim3(64), dim3(256), 0, s, c);
}

void launchNeighborSets(const Ctx& c, uint32_t scene, uint32_t count, uint32_t* counts, uint64_t* prints,
                        uint32_t* total, hipStream_t s)
{
    hipLaunchKernelGGL(kNeighborSets, dim3(gridFor(count, 256)), dim3(256), 0, s, c, scene, count, counts, prints,
                       total);
}

void launchDensityInfo(const Ctx& c, uint32_t scene, uint32_t count, float threshold, double* partials,
                       uint32_t* upartials, hipStream_t s)
{
    hipLaunchKernelGGL(kDensityInfo, dim3(kInfoBlocks), dim3(256), 0, s, c, scene, count, threshold, partials,
                       upartials);
}

// ------------------------------------------------- division self-check --
// Compares divBy(a, b, divRecip(b)) with the division a / b on `count`
// pseudo-random pairs: a = +0 or +-m * 2^ea, b = m' * 2^eb with exponents
// uniform in [-40, 39] and random 23-bit mantissas, plus, every fourth pair,
// a in [0, 2.2 b] (the kernels' d / h range); and sqrtRanged(x) with sqrtf(x)
// for x = +0 or m * 2^e, e uniform in [-96, 126]. Counts the results whose
// bits differ.
// SplitMix64 finaliser
__device__ __forceinline__ uint64_t checkMix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

__device__ __forceinline__ float rangedFloat(uint64_t r)
{
    const uint32_t e = static_cast<uint32_t>((r >> 23) % 80U) + 87U;  // 2^-40 .. 2^39
    return __uint_as_float((e << 23) | static_cast<uint32_t>(r & 0x7FFFFFU));
}

__global__ __launch_bounds__(256) void kDivisionCheck(uint64_t seed, uint64_t count, unsigned long long* bad)
{
    uint32_t local = 0;
    for (uint64_t k = blockIdx.x * 256ULL + threadIdx.x; k < count; k += static_cast<uint64_t>(gridDim.x) * 256ULL)
    {
        const uint64_t r0 = checkMix(seed ^ (k * 2ULL));
        const uint64_t r1 = checkMix(seed ^ (k * 2ULL + 1ULL));
        const float b = rangedFloat(r1);
        float a;
        if ((k & 3ULL) == 3ULL)
        {
            a = b * (2.2F * static_cast<float>(r0 >> 40) * 0x1p-24F);  // d / h range, includes +0
        }
        else if ((k & 255ULL) == 254ULL)
        {
            a = 0.0F;
        }
        else
        {
            a = rangedFloat(r0);
            a = ((r0 >> 63) != 0ULL) ? -a : a;
        }
        if (!(a == 0.0F || (fabsf(a) >= kDivLo && fabsf(a) <= kDivHi)))
        {
            continue;  // outside the domain of divBy (tiny products of the d / h case)
        }
        const float fast = divBy(a, b, divRecip(b));
        const float ieee = a / b;
        local += (__float_as_uint(fast) != __float_as_uint(ieee)) ? 1U : 0U;
        // sqrtRanged against sqrtf: x = +0 or exponents uniform in [-96, 126]
        const uint64_t r2 = checkMix(seed ^ (k * 2ULL + 0x5A5A5A5AULL));
        const uint32_t ex = static_cast<uint32_t>((r2 >> 23) % 223U) + 31U;
        const float x = (k & 255ULL) == 253ULL ? 0.0F : __uint_as_float((ex << 23) | static_cast<uint32_t>(r2 & 0x7FFFFFU));
        local += (__float_as_uint(sqrtRanged(x)) != __float_as_uint(sqrtf(x))) ? 1U : 0U;
    }
    if (local != 0U)
    {
        atomicAdd(bad, static_cast<unsigned long long>(local));
    }
}

void launchDivisionCheck(uint64_t seed, uint64_t count, unsigned long long* bad, hipStream_t s)
{
    hipLaunchKernelGGL(kDivisionCheck, dim3(4096), dim3(256), 0, s, seed, count, bad);
}

}  // namespace asph
