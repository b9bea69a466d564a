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

// ------------------------------------------------------------------ prep --
// Per step: dense offsets of the scenes, simulated time, sloshing gravity,
// refinement-zone centre and the refinement cadence
// (AdaptiveSphSimulation.cu:75-79, frame counter incremented at :84).
__global__ __launch_bounds__(256) void kPrep(Ctx c)
{
    __shared__ uint32_t lds[4];
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
    if (threadIdx.x == 0)
    {
        *c.nTotal = carry;
    }
}

// ------------------------------------------------- kick + drift + key --
// halfKickVelocities (Algorithm.cu:270-284), updatePositions (:286-298),
// assignParticlesToCells (NeighborKernels.cu:5-15) and the cell histogram.
__global__ __launch_bounds__(256) void kKickDriftKey(Ctx c)
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
        atomicAdd(&c.cellCount[key], 1U);
    }
}

// ------------------------------------------------------- scene ranges --
// Sorted range of each scene and the slot-0 omission of
// calculateCellStartAndEndIndices (NeighborKernels.cu:27-36): when the
// lowest-key cell of a scene holds exactly one particle, that cell and the
// next occupied cell are invisible to every neighbour walk.
__global__ void kSceneRanges(Ctx c)
{
    const uint32_t nOrder = *c.nOrderPtr;
    for (uint32_t k = threadIdx.x; k < nOrder; k += blockDim.x)
    {
        const uint32_t s = c.cellOrder[k];
        const SceneDev& sc = c.scenes[s];
        SceneState& st = c.states[s];
        const uint32_t ss = c.cellOff[sc.cellBase];
        c.sortedStarts[k] = ss;
        st.sortedStart = ss;
        const uint32_t n = st.count;
        uint32_t lost = ss;
        if ((sc.flags & kSceneFixCellStart) == 0U && n > 0U)
        {
            if (n == 1U)
            {
                lost = ss + 1U;
            }
            else
            {
                const uint32_t k0 = c.sortedKeys[ss];
                const uint32_t k1 = c.sortedKeys[ss + 1U];
                if (k0 != k1)
                {
                    lost = c.cellOff[k1 + 1U];
                }
            }
        }
        st.lostEnd = lost;
    }
}

// --------------------------------------------------------------- gather --
__global__ __launch_bounds__(256) void kGather(Ctx c)
{
    const uint32_t n = *c.nTotal;
    const uint32_t nOrder = *c.nOrderPtr;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    {
        const uint32_t k = segmentOf(c.sortedStarts, nOrder, t);
        const uint32_t i = c.sortedVals[t];
        const F4 p = c.pos[i];
        const F4 v = c.vel[i];
        const float hh = c.h[i];
        const float r = 2.0F * hh;
        c.s0[t] = F4 {p.x, p.y, p.z, r * r};
        c.s1[t] = F4 {v.x, v.y, v.z, hh};
        c.sm[t] = c.m[i];
        c.sscene[t] = c.cellOrder[k];
    }
}

// -------------------------------------------------------------- density --
// computeDensities (Algorithm.cu:17-54) and computeBoundaryDensityContribution
// (BoundaryAlgorithm.cu:15-51). Phase 1 walks the candidate cells in the
// reference's order (x, then y, then z innermost; slots ascending) and stores
// the accepted slots in an entry-major neighbour list; phase 2 accumulates the
// density over that list with every lane active.
__global__ __launch_bounds__(256) void kDensity(Ctx c)
{
    const uint32_t n = *c.nTotal;
    const uint32_t K = c.listCap;
    const uint32_t LS = c.pCap;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    {
        const uint32_t s = c.sscene[t];
        const SceneDev& sc = c.scenes[s];
        const uint32_t lost = c.states[s].lostEnd;
        const F4 q = c.s0[t];
        const float hi = c.s1[t].w;
        const float R = 2.0F * hi;
        const CellBox b = cellBox(sc, q.x, q.y, q.z, R);
        const int gxgy = sc.gx * sc.gy;
        uint32_t cnt = 0;
        for (int x = b.x0; x <= b.x1; ++x)
        {
            for (int y = b.y0; y <= b.y1; ++y)
            {
                for (int z = b.z0; z <= b.z1; ++z)
                {
                    const uint32_t cell = sc.cellBase + static_cast<uint32_t>(x + y * sc.gx + z * gxgy);
                    const uint32_t lo = max(c.cellOff[cell], lost);
                    const uint32_t hiS = c.cellOff[cell + 1U];
                    for (uint32_t j = lo; j < hiS; ++j)
                    {
                        const F4 o = c.s0[j];
                        const float dx = o.x - q.x;
                        const float dy = o.y - q.y;
                        const float dz = o.z - q.z;
                        const float d2 = dot3w(dx, dy, dz, dx, dy, dz);
                        if (d2 <= o.w)
                        {
                            if (cnt < K)
                            {
                                c.list[static_cast<size_t>(cnt) * LS + t] = j;
                            }
                            ++cnt;
                        }
                    }
                }
            }
        }
        float rho = 0.0F;
        if (cnt <= K)
        {
            for (uint32_t e = 0; e < cnt; ++e)
            {
                const uint32_t j = c.list[static_cast<size_t>(e) * LS + t];
                const F4 o = c.s0[j];
                const float dx = o.x - q.x;
                const float dy = o.y - q.y;
                const float dz = o.z - q.z;
                const float d2 = dot3w(dx, dy, dz, dx, dy, dz);
                const float d = sqrtf(d2);
                rho = rho + c.sm[j] * wendland(d, c.s1[j].w);
            }
        }
        else
        {
            for (int x = b.x0; x <= b.x1; ++x)
            {
                for (int y = b.y0; y <= b.y1; ++y)
                {
                    for (int z = b.z0; z <= b.z1; ++z)
                    {
                        const uint32_t cell = sc.cellBase + static_cast<uint32_t>(x + y * sc.gx + z * gxgy);
                        const uint32_t lo = max(c.cellOff[cell], lost);
                        const uint32_t hiS = c.cellOff[cell + 1U];
                        for (uint32_t j = lo; j < hiS; ++j)
                        {
                            const F4 o = c.s0[j];
                            const float dx = o.x - q.x;
                            const float dy = o.y - q.y;
                            const float dz = o.z - q.z;
                            const float d2 = dot3w(dx, dy, dz, dx, dy, dz);
                            if (d2 <= o.w)
                            {
                                rho = rho + c.sm[j] * wendland(sqrtf(d2), c.s1[j].w);
                            }
                        }
                    }
                }
            }
        }
        // boundary contribution (always added, as the reference launches it unconditionally)
        float rhoB = 0.0F;
        if (sc.bCount > 0U)
        {
            const float r2 = R * R;
            for (int x = b.x0; x <= b.x1; ++x)
            {
                for (int y = b.y0; y <= b.y1; ++y)
                {
                    for (int z = b.z0; z <= b.z1; ++z)
                    {
                        const uint32_t cell = sc.cellBase + static_cast<uint32_t>(x + y * sc.gx + z * gxgy);
                        const uint32_t lo = max(c.bcellOff[cell], sc.bLostEnd);
                        const uint32_t hiS = c.bcellOff[cell + 1U];
                        for (uint32_t j = lo; j < hiS; ++j)
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
        rho = rho + rhoB;
        const float p = tait(rho, sc);
        c.s2[t] = F4 {rho, p, p / (rho * rho), 0.0F};
        c.ncount[t] = cnt;
    }
}

// --------------------------------------------------------------- forces --
struct PairAcc
{
    float apx, apy, apz;
    float avx, avy, avz;
    float asx, asy, asz;
};

// One fluid pair of computePressureAccelerations (Algorithm.cu:72-105),
// computeViscosityAccelerations (:129-171) and
// computeSurfaceTensionAccelerations (:192-218), accumulated separately.
__device__ __forceinline__ void forcePair(const Ctx& c, uint32_t j, const F4& q, const F4& vi, float rhoi, float prri,
                                          float nuNum, float epsH2, bool doVisc, bool doST, PairAcc& A)
{
    const F4 o = c.s0[j];
    const float ox = q.x - o.x;
    const float oy = q.y - o.y;
    const float oz = q.z - o.z;
    const float d2 = dot3w(ox, oy, oz, ox, oy, oz);
    const F4 rj = c.s2[j];
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
    const F4 nj = c.s1[j];
    const float mj = c.sm[j];
    const float pterm = prri + rj.z;
    const float dw = wendlandGrad(d, nj.w);
    A.apx = A.apx - ((mj * dx) * dw) * pterm;
    A.apy = A.apy - ((mj * dy) * dw) * pterm;
    A.apz = A.apz - ((mj * dz) * dw) * pterm;
    if (doVisc)
    {
        const float vx = vi.x - nj.x;
        const float vy = vi.y - nj.y;
        const float vz = vi.z - nj.z;
        const float comp = dot3w(vx, vy, vz, ox, oy, oz);
        if (comp < 0.0F)
        {
            const float nu = nuNum / (rhoi + rj.x);
            const float piv = ((-nu) * comp) / (d2 + epsH2);
            const float sv = (mj * piv) * dw;
            A.avx = A.avx - sv * dx;
            A.avy = A.avy - sv * dy;
            A.avz = A.avz - sv * dz;
        }
    }
    if (doST)
    {
        const float sw = mj * wendland(d, nj.w);
        A.asx = A.asx + sw * dx;
        A.asy = A.asy + sw * dy;
        A.asz = A.asz + sw * dz;
    }
}

// computeExternalAccelerations + fluid pair forces + boundary pressure and
// friction (BoundaryAlgorithm.cu:53-162), added to g in the reference's launch
// order (SphSimulation.cu:110-115), then the second half kick (:117) unless the
// scene refines this step (the kick then follows the refinement).
__global__ __launch_bounds__(256) void kForce(Ctx c)
{
    const uint32_t n = *c.nTotal;
    const uint32_t K = c.listCap;
    const uint32_t LS = c.pCap;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    {
        const uint32_t s = c.sscene[t];
        const SceneDev& sc = c.scenes[s];
        const SceneState& st = c.states[s];
        const F4 q = c.s0[t];
        const F4 vi = c.s1[t];
        const float hi = vi.w;
        const F4 ri = c.s2[t];
        const float rhoi = ri.x;
        const float prri = ri.z;
        const uint32_t cnt = c.ncount[t];
        const bool doVisc = sc.alpha != 0.0F;
        const bool doST = sc.sigma != 0.0F;
        const float nuNum = ((2.0F * sc.alpha) * hi) * sc.c0;
        const float epsH2 = (0.01F * hi) * hi;
        const float R = 2.0F * hi;
        const CellBox b = cellBox(sc, q.x, q.y, q.z, R);
        const int gxgy = sc.gx * sc.gy;
        PairAcc A {0.F, 0.F, 0.F, 0.F, 0.F, 0.F, 0.F, 0.F, 0.F};
        if (cnt <= K)
        {
            for (uint32_t e = 0; e < cnt; ++e)
            {
                const uint32_t j = c.list[static_cast<size_t>(e) * LS + t];
                if (j != t)
                {
                    forcePair(c, j, q, vi, rhoi, prri, nuNum, epsH2, doVisc, doST, A);
                }
            }
        }
        else
        {
            const uint32_t lost = st.lostEnd;
            for (int x = b.x0; x <= b.x1; ++x)
            {
                for (int y = b.y0; y <= b.y1; ++y)
                {
                    for (int z = b.z0; z <= b.z1; ++z)
                    {
                        const uint32_t cell = sc.cellBase + static_cast<uint32_t>(x + y * sc.gx + z * gxgy);
                        const uint32_t lo = max(c.cellOff[cell], lost);
                        const uint32_t hiS = c.cellOff[cell + 1U];
                        for (uint32_t j = lo; j < hiS; ++j)
                        {
                            if (j == t)
                            {
                                continue;
                            }
                            const F4 o = c.s0[j];
                            const float dx = q.x - o.x;
                            const float dy = q.y - o.y;
                            const float dz = q.z - o.z;
                            const float d2 = dot3w(dx, dy, dz, dx, dy, dz);
                            if (d2 <= o.w)
                            {
                                forcePair(c, j, q, vi, rhoi, prri, nuNum, epsH2, doVisc, doST, A);
                            }
                        }
                    }
                }
            }
        }
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
        if (doST)
        {
            const float k = sc.sigma / c.sm[t];
            ax = ax + k * A.asx;
            ay = ay + k * A.asy;
            az = az + k * A.asz;
        }
        float bpx = 0.F, bpy = 0.F, bpz = 0.F;
        float bfx = 0.F, bfy = 0.F, bfz = 0.F;
        if (sc.bCount > 0U)
        {
            const float r2 = R * R;
            for (int x = b.x0; x <= b.x1; ++x)
            {
                for (int y = b.y0; y <= b.y1; ++y)
                {
                    for (int z = b.z0; z <= b.z1; ++z)
                    {
                        const uint32_t cell = sc.cellBase + static_cast<uint32_t>(x + y * sc.gx + z * gxgy);
                        const uint32_t lo = max(c.bcellOff[cell], sc.bLostEnd);
                        const uint32_t hiS = c.bcellOff[cell + 1U];
                        for (uint32_t j = lo; j < hiS; ++j)
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
                            const float dw = wendlandGrad(d, hi);
                            const float sp = (psi * prri) * dw;
                            bpx = bpx - sp * dx;
                            bpy = bpy - sp * dy;
                            bpz = bpz - sp * dz;
                            const float comp = dot3w(vi.x, vi.y, vi.z, ox, oy, oz);
                            if (comp < 0.0F)
                            {
                                const float f = c.bvisc[j];
                                const float nu = ((f * hi) * sc.c0) / (2.0F * rhoi);
                                const float piv = ((-nu) * comp) / (d2 + epsH2);
                                bfx = bfx - ((dx * psi) * piv) * dw;
                                bfy = bfy - ((dy * psi) * piv) * dw;
                                bfz = bfz - ((dz * psi) * piv) * dw;
                            }
                        }
                    }
                }
            }
        }
        ax = ax + bpx;
        ay = ay + bpy;
        az = az + bpz;
        ax = ax + bfx;
        ay = ay + bfy;
        az = az + bfz;
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

__device__ __forceinline__ bool mergeParticipant(const SceneDev& sc, const SceneState& st)
{
    return (sc.flags & kSceneActive) != 0U && st.refineNow != 0U && st.countBefore > 1U;
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
            c.ckeysA[d] = ~0ULL;
            c.cvalsA[d] = 0U;
            continue;
        }
        const uint32_t l = d - st.denseOff;
        c.cand[d] = 0U;  // reset value of mergeCandidates (AdaptiveSphSimulation.cu:127)
        c.removal[d] = 0U;
        c.keep[d] = 0U;
        const float v = mergeCriterion(c, sc, st, sc.pOff + l);
        c.ckeysA[d] = (static_cast<uint64_t>(s) << 32U) | static_cast<uint64_t>(orderedBits(v));
        c.cvalsA[d] = l;
        if (v > 0.0F)
        {
            atomicAdd(&st.positives, 1U);
        }
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
    const uint32_t n = *c.nTotal;
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
        if (e >= st.mergeK)
        {
            continue;
        }
        const uint32_t id = c.csortedVals[t];
        const uint32_t i = sc.pOff + id;
        const F4 q = c.pos[i];
        const float hh = c.h[i];
        const float massLimit = sc.maxMassRatio * sc.m0;
        float closest = FLT_MAX;
        uint32_t best = kNoCandidate;
        walkCanonical(c, sc, st.lostEnd, q.x, q.y, q.z, hh, [&](uint32_t j) {
            if (j == i)
            {
                return;
            }
            const F4 pj = c.pos[j];
            const float ox = pj.x - q.x;
            const float oy = pj.y - q.y;
            const float oz = pj.z - q.z;
            const float d2 = dot3w(ox, oy, oz, ox, oy, oz);
            const float hmin = gmin(hh, c.h[j]);
            if (d2 < hmin * hmin && d2 < closest && c.m[j] < massLimit)
            {
                closest = d2;
                best = j - sc.pOff;
            }
        });
        c.cand[st.denseOff + id] = best;
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
        if (c.cand[st.denseOff + cc] == l && l < cc)
        {
            c.keep[d] = 1U;
            c.removal[st.denseOff + cc] = 1U;
            atomicAdd(&st.mergePairs, 1U);
        }
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
        c.h[a] = detCbrt((kMergeCoef * M) / denom);
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
    }
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
            c.ckeysA[d] = ~0ULL;
            c.cvalsA[d] = 0U;
            continue;
        }
        const float v = splitCriterion(c, sc, st, sc.pOff + l);
        c.ckeysA[d] = (static_cast<uint64_t>(s) << 32U) | static_cast<uint64_t>(~orderedBits(v));
        c.cvalsA[d] = l;
        if (v > 0.0F)
        {
            atomicAdd(&st.positives, 1U);
        }
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
    const uint32_t n = *c.nTotal;
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
    hipLaunchKernelGGL(kKickDriftKey, dim3(gridFor(c.pCap, lc.blockStream)), dim3(lc.blockStream), 0, s, c);
}

void launchSceneRanges(const Ctx& c, hipStream_t s) { hipLaunchKernelGGL(kSceneRanges, dim3(1), dim3(256), 0, s, c); }

void launchGather(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kGather, dim3(gridFor(c.pCap, lc.blockStream)), dim3(lc.blockStream), 0, s, c);
}

void launchDensity(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kDensity, dim3(gridFor(c.pCap, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchForce(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kForce, dim3(gridFor(c.pCap, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchMergeCriteria(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kMergeCriteria, dim3(gridFor(c.pCap, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchMergeSelect(const Ctx& c, hipStream_t s) { hipLaunchKernelGGL(kMergeSelect, dim3(1), dim3(64), 0, s, c); }

void launchMergeCandidates(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kMergeCandidates, dim3(gridFor(c.pCap, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchResolvePairs(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kResolvePairs, dim3(gridFor(c.pCap, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchMergeApply(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kMergeApply, dim3(gridFor(c.pCap, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchCompact(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kCompact, dim3(gridFor(c.pCap, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchAfterMerge(const Ctx& c, hipStream_t s) { hipLaunchKernelGGL(kAfterMerge, dim3(1), dim3(64), 0, s, c); }

void launchCopyBack(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kCopyBack, dim3(gridFor(c.pCap, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchSplitCriteria(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kSplitCriteria, dim3(gridFor(c.pCap, 256)), dim3(256), 0, s, c);
    (void)lc;
}

void launchSplitSelect(const Ctx& c, hipStream_t s) { hipLaunchKernelGGL(kSplitSelect, dim3(1), dim3(64), 0, s, c); }

void launchSplitApply(const Ctx& c, const LaunchCfg& lc, hipStream_t s)
{
    hipLaunchKernelGGL(kSplitApply, dim3(gridFor(c.pCap, 256)), dim3(256), 0, s, c);
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

}  // namespace asph
