// Shared definitions of the HIP engine: the scene table, per-scene mutable
// state, and the device math of the SPH pipeline.
//
// Arithmetic discipline (every file is compiled with -ffp-contract=off and
// IEEE-correct FP32 division / square root): each expression follows the
// reference statement it restates in the same association order, so cell
// keys, neighbour sets, densities, forces and trajectories are bit-identical
// to the host restatement run in FP32. Quotients that share a divisor go
// through one reciprocal (divRecip / divBy) where that yields the quotient's
// exact bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace asph
{

constexpr uint32_t kNoCandidate = 0xFFFFFFFFU;  // refinement/Common.cuh:119-122
constexpr uint32_t kWave = 64;
#ifndef ASPH_GROUP
#define ASPH_GROUP 16
#endif
constexpr uint32_t kGroup = ASPH_GROUP;  // sorted slots per culling box (a power of two, at most 64)
constexpr uint32_t kBoundaryGroup = 16;  // boundary slots per culling box (chunks of a cell's range)

// Query order modes (Ctx::queryTileBits, with qperm): 1 and 2 Morton 2x2x2 /
// 4x4x4 sub-cells, 3 x-major 4x4x4 slabs, kQueryAuto: x-major slabs for the
// scenes whose waves have loosened (SceneState::qorder, set by kPrep from the
// list builder's counts), the sorted order for the others.
constexpr uint32_t kQueryAuto = 4;
constexpr uint32_t kQuerySample = 64;       // one wave in 64 counts for the automatic order
constexpr uint32_t kQueryMinHits = 65536;   // neighbours counted before a decision (~16 waves)

// Wave order of the list builder: waves whose last step cost more than 1.5x
// the sampled mean survivors, or walked lane by lane, run first (at most
// kHeavyCap of them), so that no costly wave starts in the kernel's last
// round and finishes alone (1M dam break over 1,000 steps, ms/step by list
// size: none 0.997, 256 0.985, 1,024 0.960, 4,096 0.960).
#ifndef ASPH_HEAVY_CAP
#define ASPH_HEAVY_CAP 1024
#endif
#ifndef ASPH_HEAVY_FACTOR2
#define ASPH_HEAVY_FACTOR2 3
#endif
constexpr uint32_t kHeavyCap = ASPH_HEAVY_CAP;
constexpr uint32_t kHeavyFactor2 = ASPH_HEAVY_FACTOR2;  // threshold: kHeavyFactor2 / 2 x the mean cost
constexpr uint32_t kCostPerLane = 0xFFFFFFFFU;  // wave cost of a lane-by-lane walk
constexpr uint32_t kCostSample = 16;            // one wave in 16 feeds the mean

// Neighbour-list row width: each slot's entries are stored kListRowEntries
// at a time (ASPH_LIST_ROW_CHUNKS 16-byte chunks side by side); list
// capacities are rounded up to a multiple of it. Two chunks per row fill
// whole 32-byte sectors at each write (1M dam break: list-builder writes
// 628 -> 492 MB per step), but each 16-byte list load then spans twice the
// cache lines and the tolerance-mode force kernel ran 6 % slower, so one.
#ifndef ASPH_LIST_ROW_CHUNKS
#define ASPH_LIST_ROW_CHUNKS 1
#endif
constexpr uint32_t kListRowChunks = ASPH_LIST_ROW_CHUNKS;
constexpr uint32_t kListRowEntries = 4U * kListRowChunks;

// kNeighbors deals its regular waves to the XCDs (each with its own L2) in
// runs of kNeighborRunWaves consecutive query waves, so that the waves that
// read the same candidate cells share an L2 while costly regions still
// spread over every XCD (0: the dispatch order, workgroup i on XCD i mod 8).
// 1M dam break, ms/step over 200 steps by run length in workgroups of 4
// waves: none 0.905, 16: 0.904, 64: 0.895, 256: 0.884, one contiguous eighth
// per XCD 0.891 (the 256-scene batch 4 % slower with eighths); list-builder
// reads (FETCH_SIZE) 770 -> 648 MB at 256.
#ifndef ASPH_NEIGHBOR_RUN_WAVES
#define ASPH_NEIGHBOR_RUN_WAVES 1024
#endif

// Waves per SIMD that kNeighbors' register budget is fitted to (80 VGPRs at
// 6). 1M dam break, ms/step over 200 steps: 6: 0.883, 7 (72 VGPRs, 32 B of
// spills): 0.891, 8 (64 VGPRs, 100 B of spills): 0.912.
#ifndef ASPH_NEIGHBOR_WAVES
#define ASPH_NEIGHBOR_WAVES 6
#endif
constexpr uint32_t kNeighborRunWaves = ASPH_NEIGHBOR_RUN_WAVES;

// Scene flags (SceneDev::flags)
constexpr uint32_t kSceneActive = 1U;
constexpr uint32_t kSceneAdaptive = 2U;
constexpr uint32_t kSceneFixCellStart = 4U;
constexpr uint32_t kSceneBoundaryInside = 8U;  // some boundary particle lies in an interior cell
constexpr uint32_t kSceneKernelMergedH = 16U;  // merged h with the kernel normalisation
constexpr uint32_t kSceneUniformHM = 32U;      // every particle has h = h0 and m = m0 (plain scenes)
constexpr uint32_t kSceneToleranceMath = 64U;  // density and force sums in the tolerance mode

// Read-only per-scene description (uploaded when the scene set changes).
struct SceneDev
{
    float dmin[3];
    float dmax[3];
    float friction;
    float g[3];
    float c0, rho0, alpha, sigma, vmax, h0, r0, m0, gamma;
    float B;  // ((rho0 * c0) * c0) / gamma (Common.cuh:9)
    float nuLam;  // kinematic viscosity of the Morris laminar term (0 = off)
    float sloshA, sloshT;
    float dt, halfDt;
    int gx, gy, gz;
    float cs[3];
    uint32_t cellBase, nCells;  // global cell range [cellBase, cellBase + nCells] (+1 end entry)
    uint32_t pOff, pCap;        // canonical particle region
    uint32_t bOff, bCount;      // boundary particles (sorted by cell)
    uint32_t bLostEnd;          // boundary slots [bOff, bLostEnd) omitted (slot-0 quirk)
    uint32_t flags;
    uint32_t tpb;
    // refinement (RefinementParameters.cuh:80-104 + zone extension)
    uint32_t criterion, initialCooldown, cooldown, maxCount;
    float minMassRatio, maxMassRatio, batchRatio;
    float eps, alphaSplit, centerRatio, vertexRatio;
    float velSplit, velMerge, vortSplit, vortMerge, ifSplit, ifMerge;
    float zoneC[3], zoneR, zonePathR, zonePathT;
};

// Mutable per-scene state, owned by the device.
struct SceneState
{
    uint32_t count;        // current fluid particle count
    uint32_t frame;        // refinement frame counter (AdaptiveSphSimulation.cuh:43)
    double time;           // simulated time at the end of the last step
    uint32_t refineNow;    // 1 when this step refines
    float gEff[3];         // gravity of this step (sloshing extension)
    float zoneNow[3];      // refinement-zone centre of this step
    uint32_t denseOff;     // exclusive prefix of counts over scenes (step start)
    uint32_t sortedStart;  // first sorted slot of this scene
    uint32_t lostEnd;      // fluid slots [sortedStart, lostEnd) omitted (slot-0 quirk)
    // refinement bookkeeping of the last refinement
    uint32_t countBefore;
    uint32_t positives;
    uint32_t segStart;
    uint32_t mergeK;
    uint32_t mergePairs;
    uint32_t removed;
    uint32_t countAfterMerge;
    uint32_t splitK;
    uint32_t splitDone;
    uint32_t countAfter;
    uint32_t refined;
    // automatic query order (query tile mode 4): on once this scene's waves
    // have loosened (slabbed queries from then on); the list builder's
    // survivor and neighbour counts of the last step decide
    uint32_t qorder;
    uint32_t qsurv;
    uint32_t qhits;
};

struct F4
{
    float x, y, z, w;
};

// Scenes whose fluid pair forces take the tolerance loop (kForceTol): the
// tolerance mode without surface tension or the Morris term.
__host__ __device__ inline bool toleranceForce(const SceneDev& sc)
{
    return (sc.flags & kSceneToleranceMath) != 0U && sc.sigma == 0.0F && sc.nuLam == 0.0F;
}

// ------------------------------------------------------------- math --

__device__ __forceinline__ float gmin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float gmax(float a, float b) { return (a < b) ? b : a; }

// glm vec4 dot with w = 0: (x*x' + y*y') + (z*z' + 0)
__device__ __forceinline__ float dot3w(float ax, float ay, float az, float bx, float by, float bz)
{
    const float tx = ax * bx;
    const float ty = ay * by;
    const float tz = az * bz;
    return (tx + ty) + (tz + 0.0F);
}

// glm vec4 self dot: (x*x + y*y) + (z*z + w*w)
__device__ __forceinline__ float dot4(float x, float y, float z, float w)
{
    return (x * x + y * y) + (z * z + w * w);
}

// glm vec3 dot: (x*x' + y*y') + z*z'
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    const float tx = ax * bx;
    const float ty = ay * by;
    const float tz = az * bz;
    return (tx + ty) + tz;
}

constexpr float kPi = 3.141592653589793F;               // WendlandKernel.cuh:10
constexpr float kWCoef = 21.F / (16.F * kPi);            // WendlandKernel.cuh:11
constexpr float kWDCoef = -5.F * kWCoef;                 // WendlandKernel.cuh:13
constexpr float kMergeCoef = 16.0F * 3.14159265358979323846F / 21.F;  // AdaptiveAlgorithm.cu:64

// wendlandKernel (WendlandKernel.cuh:39-52)
__device__ __forceinline__ float wendland(float d, float h)
{
    const float q = d / h;
    if (q > 2.0F)
    {
        return 0.0F;
    }
    const float h3 = (h * h) * h;
    const float tmp = 1.0F - 0.5F * q;
    const float tmp4 = (tmp * tmp) * (tmp * tmp);
    return ((kWCoef / h3) * tmp4) * (2.0F * q + 1.0F);
}

// wendlandDerivativeKernel (WendlandKernel.cuh:69-84)
__device__ __forceinline__ float wendlandGrad(float d, float h)
{
    const float q = d / h;
    if (q > 2.0F)
    {
        return 0.0F;
    }
    const float h3 = (h * h) * h;
    const float h4 = h3 * h;
    const float tmp = 1.0F - 0.5F * q;
    const float tmp3 = (tmp * tmp) * tmp;
    return ((kWDCoef / h4) * q) * tmp3;
}

// ---- division by a shared divisor, same bits as the IEEE quotient ----------
// The correctly rounded FP32 division a / b that hipcc emits for gfx950 is:
// range scaling of a and b (v_div_scale), r = rcp(b), e = fma(-b, r, 1),
// y = fma(e, r, r), q0 = a * y, r0 = fma(-b, q0, a), q1 = fma(r0, y, q0),
// r1 = fma(-b, q1, a), fma(r1, y, q1) with a scale fix-up (v_div_fmas), and a
// special-value fix-up (v_div_fixup). When a is +0 or |a|, |b| lie in
// [2^-40, 2^40] the scalings are the identity (no operand or quotient near
// the denormal or overflow range), v_div_fmas is a plain fma and v_div_fixup
// returns its input, so divBy gives the quotient's exact bits while y, which
// depends on b only, is shared by every numerator over b.
constexpr float kDivLo = 0x1p-40F;
constexpr float kDivHi = 0x1p40F;

__device__ __forceinline__ float divRecip(float b)
{
    const float r = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, r, 1.0F);
    return __builtin_fmaf(e, r, r);
}

__device__ __forceinline__ float divBy(float a, float b, float y)
{
    const float q0 = a * y;
    const float r0 = __builtin_fmaf(-b, q0, a);
    const float q1 = __builtin_fmaf(r0, y, q0);
    const float r1 = __builtin_fmaf(-b, q1, a);
    return __builtin_fmaf(r1, y, q1);
}

// Correctly rounded FP32 square root as hipcc emits it for gfx950: inputs
// below 2^-96 scaled by 2^32, s = v_sqrt(x), then s - 1 ulp or s + 1 ulp
// chosen by the signs of fma(-(s -+ ulp), s, x), the result unscaled by
// 2^-16, and zeros and infinities passed through. For x = +0 or x in
// [2^-96, 2^127) the scalings and the pass-through are identities (+0 runs
// through the corrections unchanged), so the steps below give sqrtf's bits.
constexpr float kSqrtLo = 0x1p-96F;
constexpr float kSqrtHi = 0x1p127F;

__device__ __forceinline__ float sqrtRanged(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1U);
    const float sp = __uint_as_float(__float_as_uint(s) + 1U);
    float t = (__builtin_fmaf(-sm, s, x) <= 0.0F) ? sm : s;
    t = (__builtin_fmaf(-sp, s, x) > 0.0F) ? sp : t;
    return t;
}

// x in the domain of sqrtRanged
__device__ __forceinline__ bool sqrtInRange(float x) { return x == 0.0F || (x >= kSqrtLo && x < kSqrtHi); }

// wendlandKernel / wendlandDerivativeKernel (WendlandKernel.cuh:39-52, 69-84)
// given q = d / h and the per-particle factor; wendlandC and wendlandGradC
// below are these with q = d / h
__device__ __forceinline__ float wendlandQ(float q, float coefOverH3)
{
    // evaluated for every q and selected (no branch around the terms)
    const float tmp = 1.0F - 0.5F * q;
    const float tmp4 = (tmp * tmp) * (tmp * tmp);
    const float w = (coefOverH3 * tmp4) * (2.0F * q + 1.0F);
    return q > 2.0F ? 0.0F : w;
}

__device__ __forceinline__ float wendlandGradQ(float q, float dcoefOverH4)
{
    if (q > 2.0F)
    {
        return 0.0F;
    }
    const float tmp = 1.0F - 0.5F * q;
    const float tmp3 = (tmp * tmp) * tmp;
    return (dcoefOverH4 * q) * tmp3;
}

// wendland with (coef / h^3) precomputed per particle (same rounding)
__device__ __forceinline__ float wendlandC(float d, float h, float coefOverH3)
{
    return wendlandQ(d / h, coefOverH3);
}

// wendlandGrad with (dcoef / h^4) precomputed per particle
__device__ __forceinline__ float wendlandGradC(float d, float h, float dcoefOverH4)
{
    return wendlandGradQ(d / h, dcoefOverH4);
}

__device__ __forceinline__ float coefOverH3(float h) { return kWCoef / ((h * h) * h); }
__device__ __forceinline__ float dcoefOverH4(float h) { return kWDCoef / (((h * h) * h) * h); }

// powf(x, g) of computeTaitPressure: integral g by binary exponentiation in
// FP64 rounded to FP32; otherwise FP64 pow rounded to FP32.
__device__ __forceinline__ float detPow(float x, float g)
{
    const bool integral = (g == floorf(g)) && g >= 1.0F && g <= 32.0F;
    if (integral)
    {
        double base = static_cast<double>(x);
        double result = 1.0;
        int e = static_cast<int>(g);
        while (e != 0)
        {
            if ((e & 1) != 0)
            {
                result = result * base;
            }
            e >>= 1;
            if (e != 0)
            {
                base = base * base;
            }
        }
        return static_cast<float>(result);
    }
    return static_cast<float>(pow(static_cast<double>(x), static_cast<double>(g)));
}

// computeTaitPressure (Common.cuh:6-12)
__device__ __forceinline__ float tait(float density, const SceneDev& sc)
{
    const float ratio = density / sc.rho0;
    return sc.B * (detPow(ratio, sc.gamma) - 1.F);
}

// cbrtf: FP64 Newton on the mantissa from a fixed start, rounded to FP32.
__device__ __forceinline__ float detCbrt(float xf)
{
    double x = xf;
    if (x == 0.0 || !isfinite(x))
    {
        return xf;
    }
    const bool neg = x < 0.0;
    if (neg)
    {
        x = -x;
    }
    int e = 0;
    double m = frexp(x, &e);
    int q = 0;
    if (e >= 0)
    {
        q = e / 3;
    }
    else
    {
        q = -((-e + 2) / 3);
    }
    const int r = e - 3 * q;
    m = ldexp(m, r);
    double y = 0.7 + 0.23 * m;
    for (int it = 0; it < 7; ++it)
    {
        const double y2 = y * y;
        y = y - (y * y2 - m) / (3.0 * y2);
    }
    double res = ldexp(y, q);
    if (neg)
    {
        res = -res;
    }
    return static_cast<float>(res);
}

// sin / cos in FP64: Cody-Waite reduction by pi/2 and Taylor polynomials.
__device__ __forceinline__ void detSinCos(double a, double* s, double* c)
{
    constexpr double pio2_hi = 1.5707963267341256e+00;
    constexpr double pio2_lo = 6.0771005065061922e-11;
    constexpr double two_over_pi = 6.3661977236758138e-01;
    const double k = rint(a * two_over_pi);
    const double r = (a - k * pio2_hi) - k * pio2_lo;
    const double r2 = r * r;
    double ps = 1.0 / 355687428096000.0;
    ps = ps * r2 * -1.0 + 1.0 / 1307674368000.0;
    ps = ps * r2 * -1.0 + 1.0 / 6227020800.0;
    ps = ps * r2 * -1.0 + 1.0 / 39916800.0;
    ps = ps * r2 * -1.0 + 1.0 / 362880.0;
    ps = ps * r2 * -1.0 + 1.0 / 5040.0;
    ps = ps * r2 * -1.0 + 1.0 / 120.0;
    ps = ps * r2 * -1.0 + 1.0 / 6.0;
    ps = ps * r2 * -1.0 + 1.0;
    const double sr = ps * r;
    double pc = 1.0 / 6402373705728000.0;
    pc = pc * r2 * -1.0 + 1.0 / 20922789888000.0;
    pc = pc * r2 * -1.0 + 1.0 / 87178291200.0;
    pc = pc * r2 * -1.0 + 1.0 / 479001600.0;
    pc = pc * r2 * -1.0 + 1.0 / 3628800.0;
    pc = pc * r2 * -1.0 + 1.0 / 40320.0;
    pc = pc * r2 * -1.0 + 1.0 / 720.0;
    pc = pc * r2 * -1.0 + 1.0 / 24.0;
    pc = pc * r2 * -1.0 + 1.0 / 2.0;
    pc = pc * r2 * -1.0 + 1.0;
    const double cr = pc;
    const long long quadrant = static_cast<long long>(k) & 3LL;
    if (quadrant == 0)
    {
        *s = sr;
        *c = cr;
    }
    else if (quadrant == 1)
    {
        *s = cr;
        *c = -sr;
    }
    else if (quadrant == 2)
    {
        *s = -sr;
        *c = -cr;
    }
    else
    {
        *s = -cr;
        *c = sr;
    }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30U)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27U)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31U);
}

// Orderable unsigned image of an FP32 value (ascending order of values).
__device__ __forceinline__ uint32_t orderedBits(float v)
{
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000U) != 0U ? ~b : (b | 0x80000000U);
}

// calculateCellIndex + flattenCellIndex (NeighborGrid.cuh:166-177), local key
__device__ __forceinline__ uint32_t cellKey(const SceneDev& sc, float px, float py, float pz)
{
    const float scx = sc.dmax[0] - sc.dmin[0];
    const float scy = sc.dmax[1] - sc.dmin[1];
    const float scz = sc.dmax[2] - sc.dmin[2];
    const float rx = gmin(gmax(px - sc.dmin[0], 0.0F), scx);
    const float ry = gmin(gmax(py - sc.dmin[1], 0.0F), scy);
    const float rz = gmin(gmax(pz - sc.dmin[2], 0.0F), scz);
    const int cx = min(static_cast<int>(rx / sc.cs[0]), sc.gx - 1);
    const int cy = min(static_cast<int>(ry / sc.cs[1]), sc.gy - 1);
    const int cz = min(static_cast<int>(rz / sc.cs[2]), sc.gz - 1);
    return static_cast<uint32_t>(cx + (cy * sc.gx) + (cz * sc.gx * sc.gy));
}

// Candidate cell box of forEachFluidNeighbor (NeighborGrid.cuh:96-100)
struct CellBox
{
    int x0, y0, z0, x1, y1, z1;
};

__device__ __forceinline__ CellBox cellBox(const SceneDev& sc, float px, float py, float pz, float R)
{
    CellBox b;
    b.x0 = max(static_cast<int>(((px - sc.dmin[0]) - R) / sc.cs[0]), 0);
    b.y0 = max(static_cast<int>(((py - sc.dmin[1]) - R) / sc.cs[1]), 0);
    b.z0 = max(static_cast<int>(((pz - sc.dmin[2]) - R) / sc.cs[2]), 0);
    b.x1 = min(static_cast<int>(((px - sc.dmin[0]) + R) / sc.cs[0]), sc.gx - 1);
    b.y1 = min(static_cast<int>(((py - sc.dmin[1]) + R) / sc.cs[1]), sc.gy - 1);
    b.z1 = min(static_cast<int>(((pz - sc.dmin[2]) + R) / sc.cs[2]), sc.gz - 1);
    return b;
}

// Last index k with starts[k] <= x (starts ascending, starts[0] <= x).
__device__ __forceinline__ uint32_t segmentOf(const uint32_t* starts, uint32_t n, uint32_t x)
{
    uint32_t lo = 0;
    uint32_t hi = n;  // invariant: starts[lo] <= x, answer in [lo, hi)
    while (hi - lo > 1U)
    {
        const uint32_t mid = (lo + hi) >> 1U;
        if (starts[mid] <= x)
        {
            lo = mid;
        }
        else
        {
            hi = mid;
        }
    }
    return lo;
}

}  // namespace asph
