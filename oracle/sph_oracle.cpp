// ORACLE - TEST INFRASTRUCTURE ONLY.
//
// Host C++ restatement of the reference's per-step SPH particle pipeline
// (a-soltysik/AdaptiveSph, cuda/src), used as the correctness checker by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. It is never
// linked into, loaded by or called from the product library.
//
// Every stage follows the cited reference statement by statement, in the
// reference's association order (glm 1.0.1 dot: (x*x'+y*y')+(z*z'+w*w');
// vec3 dot: (x*x'+y*y')+z*z'), with no FMA contraction (-ffp-contract=off) and
// IEEE-correct division and square root. Three library functions whose bits
// the reference leaves to CUDA's libdevice are evaluated by deterministic
// algorithms written out below (the HIP kernels implement the same ones):
//   powf(x, g)  -> g integral: x^g by binary exponentiation in FP64, rounded
//                  to FP32 (otherwise FP64 pow rounded to FP32);
//   cbrtf(x)    -> FP64 Newton iteration from a fixed start, rounded to FP32;
//   sin / cos   -> FP64 Cody-Waite reduction + Taylor polynomials (only for
//                  the sloshing / moving-zone extensions).
// All three are within 1 ulp of the correctly rounded result (pinned against
// libm in tests/test_oracle_math.py); CUDA's own ulp pattern is unpinned.
//
// Parity pins (tests/test_oracle_math.py, test_oracle_physics.py,
// test_oracle_pins.py): closed forms of the Wendland kernel and Tait EOS, the
// interior-lattice density 1.009500783785*rho0 with 81 neighbours, momentum
// conservation, two-particle closed forms of the pressure, artificial
// viscosity and surface tension terms with h_i != h_j, one fluid particle
// against one wall particle for the boundary density, pressure and friction,
// the wall lattice and psi against an FP32 replay and an O(N^2) sum, every
// criterion at hand-chosen points and threshold edges, the vorticity of a
// rigid rotation on a lattice, merge pairing (exact tie, non-mutual chain,
// index-0 reset aliasing), mass-weighted merge, stable compaction, split
// count rounding and capacity clamp, getDensityInfo on a hand-built array,
// KDK free fall, brute-force keys / cell ranges / neighbour sets, stable
// top-k with ties and sentinels. tests/test_oracle_mutants.py checks that 20
// plausible slips planted in copies of this file each fail one of them.
//
// Template parameter T = float is the reference arithmetic; T = double is the
// FP64 oracle used to state the intrinsic FP32 noise floor.

#include <omp.h>
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../include/asph.h"
#include "../include/asph_debug.h"

namespace oracle
{

// ---------------------------------------------------------------- helpers --

template <class T>
struct V4
{
    T x, y, z, w;
};

template <class T>
struct V3
{
    T x, y, z;
};

// glm 1.0.1 compute_dot<vec4>: tmp = a*b; (tmp.x + tmp.y) + (tmp.z + tmp.w)
template <class T>
inline T dot4(const V4<T>& a, const V4<T>& b)
{
    const T tx = a.x * b.x;
    const T ty = a.y * b.y;
    const T tz = a.z * b.z;
    const T tw = a.w * b.w;
    return (tx + ty) + (tz + tw);
}

// glm 1.0.1 compute_dot<vec3>: tmp.x + tmp.y + tmp.z
template <class T>
inline T dot3(const V3<T>& a, const V3<T>& b)
{
    const T tx = a.x * b.x;
    const T ty = a.y * b.y;
    const T tz = a.z * b.z;
    return (tx + ty) + tz;
}

template <class T>
inline V4<T> sub4(const V4<T>& a, const V4<T>& b)
{
    return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w};
}

// glm::min(a, b) = (b < a) ? b : a ; glm::max(a, b) = (a < b) ? b : a
template <class T>
inline T gmin(T a, T b)
{
    return (b < a) ? b : a;
}
template <class T>
inline T gmax(T a, T b)
{
    return (a < b) ? b : a;
}

// Binary exponentiation in FP64 (deterministic IEEE multiplies).
inline double pow_int_f64(double base, int e)
{
    double result = 1.0;
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
    return result;
}

// powf(x, g) as used by computeTaitPressure (Common.cuh:6-12).
inline float det_powf(float x, float g)
{
    const bool integral = (g == std::floor(g)) && g >= 1.0F && g <= 32.0F;
    if (integral)
    {
        return static_cast<float>(pow_int_f64(static_cast<double>(x), static_cast<int>(g)));
    }
    return static_cast<float>(std::pow(static_cast<double>(x), static_cast<double>(g)));
}

// cbrtf(x): FP64 Newton on the mantissa, rounded to FP32.
inline float det_cbrtf(float xf)
{
    double x = xf;
    if (x == 0.0 || !std::isfinite(x))
    {
        return xf;
    }
    const bool neg = x < 0.0;
    if (neg)
    {
        x = -x;
    }
    int e = 0;
    double m = std::frexp(x, &e);  // x = m * 2^e, m in [0.5, 1)
    int q = 0;
    if (e >= 0)
    {
        q = e / 3;
    }
    else
    {
        q = -((-e + 2) / 3);
    }
    const int r = e - 3 * q;  // 0, 1, 2
    m = std::ldexp(m, r);     // m in [0.5, 4)
    double y = 0.7 + 0.23 * m;
    for (int it = 0; it < 7; ++it)
    {
        const double y2 = y * y;
        y = y - (y * y2 - m) / (3.0 * y2);
    }
    double res = std::ldexp(y, q);
    if (neg)
    {
        res = -res;
    }
    return static_cast<float>(res);
}

// sin / cos in FP64: reduction by pi/2 (Cody-Waite, two-term) and Taylor
// polynomials on [-pi/4, pi/4].
inline void det_sincos(double a, double* s, double* c)
{
    constexpr double pio2_hi = 1.5707963267341256e+00;
    constexpr double pio2_lo = 6.0771005065061922e-11;
    constexpr double two_over_pi = 6.3661977236758138e-01;
    const double k = std::nearbyint(a * two_over_pi);
    const double r = (a - k * pio2_hi) - k * pio2_lo;
    const double r2 = r * r;
    // sin: r - r^3/3! + ... + r^17/17!
    double ps = 1.0 / 355687428096000.0;  // 1/17!
    ps = ps * r2 * -1.0 + 1.0 / 1307674368000.0;  // -1/17!*r2 + 1/15!
    ps = ps * r2 * -1.0 + 1.0 / 6227020800.0;
    ps = ps * r2 * -1.0 + 1.0 / 39916800.0;
    ps = ps * r2 * -1.0 + 1.0 / 362880.0;
    ps = ps * r2 * -1.0 + 1.0 / 5040.0;
    ps = ps * r2 * -1.0 + 1.0 / 120.0;
    ps = ps * r2 * -1.0 + 1.0 / 6.0;
    ps = ps * r2 * -1.0 + 1.0;
    const double sr = ps * r;
    double pc = 1.0 / 6402373705728000.0;  // 1/18!
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
    switch (quadrant)
    {
    case 0:
        *s = sr;
        *c = cr;
        break;
    case 1:
        *s = cr;
        *c = -sr;
        break;
    case 2:
        *s = -sr;
        *c = -cr;
        break;
    default:
        *s = -cr;
        *c = sr;
        break;
    }
}

inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30U)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27U)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31U);
}

// ------------------------------------------------ Wendland C2 (WendlandKernel.cuh)

template <class T>
struct Const;

template <>
struct Const<float>
{
    static constexpr float pi = 3.141592653589793F;                // WendlandKernel.cuh:10
    static constexpr float wcoef = 21.F / (16.F * pi);              // :11
    static constexpr float wdcoef = -5.F * wcoef;                   // :13
    static constexpr float mergeCoef = 16.0F * 3.14159265358979323846F / 21.F;  // AdaptiveAlgorithm.cu:64
    static float pw(float x, float g) { return det_powf(x, g); }
    static float cbrt(float x) { return det_cbrtf(x); }
};

template <>
struct Const<double>
{
    static constexpr double pi = 3.14159265358979323846;
    static constexpr double wcoef = 21.0 / (16.0 * pi);
    static constexpr double wdcoef = -5.0 * wcoef;
    static constexpr double mergeCoef = 16.0 * pi / 21.0;
    static double pw(double x, double g) { return std::pow(x, g); }
    static double cbrt(double x) { return std::cbrt(x); }
};

// wendlandKernel (WendlandKernel.cuh:39-52)
template <class T>
inline T wendland(T distance, T h)
{
    const T q = distance / h;
    if (q > T(2))
    {
        return T(0);
    }
    const T h3 = (h * h) * h;
    const T tmp = T(1) - T(0.5) * q;
    const T tmp4 = (tmp * tmp) * (tmp * tmp);
    return ((Const<T>::wcoef / h3) * tmp4) * (T(2) * q + T(1));
}

// wendlandDerivativeKernel (WendlandKernel.cuh:69-84)
template <class T>
inline T wendlandDerivative(T distance, T h)
{
    const T q = distance / h;
    if (q > T(2))
    {
        return T(0);
    }
    const T h3 = (h * h) * h;
    const T h4 = h3 * h;
    const T tmp = T(1) - T(0.5) * q;
    const T tmp3 = (tmp * tmp) * tmp;
    return ((Const<T>::wdcoef / h4) * q) * tmp3;
}

// computeTaitPressure (Common.cuh:6-12); gamma is the extension parameter
template <class T>
inline T taitPressure(T density, T rho0, T c, T gamma)
{
    const T B = ((rho0 * c) * c) / gamma;
    const T ratio = density / rho0;
    return B * (Const<T>::pw(ratio, gamma) - T(1));
}

// ---------------------------------------------------------------- state --

template <class T>
struct Params
{
    V3<T> dmin, dmax;
    T friction;
    V3<T> g;
    T c0, rho0, restitution, alpha, sigma, vmax, h0, r0, m0;
    uint32_t tpb;
    T gamma;
    T sloshA, sloshT;
    uint32_t flags;
    T nuLam;  // extension: Morris laminar viscosity (0 = off)
};

template <class T>
struct Refinement
{
    bool enabled = false;
    T minMassRatio, maxMassRatio;
    uint32_t maxCount;
    T batchRatio;
    uint32_t initialCooldown, cooldown, criterion;
    T eps, alphaSplit, centerRatio, vertexRatio;
    T velSplit, velMerge, vortSplit, vortMerge, ifSplit, ifMerge;
    V3<T> zoneCenter;
    T zoneRadius, zonePathRadius, zonePathPeriod;
};

struct Grid
{
    std::vector<int32_t> keys;  // sorted cell keys (particleGridIndices)
    std::vector<int32_t> idx;   // sorted indices (particleArrayIndices)
    std::vector<int32_t> cellStart, cellEnd;
    uint32_t count = 0;
};

struct Trace
{
    asph_refinement_trace t {};
    std::vector<uint32_t> mergeList, candidates, pairKeep, splitParents;
};

template <class T>
class Sim
{
public:
    Params<T> p {};
    Refinement<T> ref {};
    uint32_t n = 0, cap = 0, bcap = 0;
    std::vector<V4<T>> pos, vel, acc;
    std::vector<T> rho, rad, h, m;
    uint32_t nb = 0;
    std::vector<V4<T>> bpos;
    std::vector<T> bpsi, bvisc, brad;
    int gx = 0, gy = 0, gz = 0;
    V3<T> cs {};
    Grid fluid, boundary;
    uint32_t frame = 0;
    double time = 0.0;
    Trace trace;

    // ------------------------------------------------------------ grid --
    void createGrid()
    {
        // NeighborGrid::createGrid / calculateGridSize (NeighborGrid.cu:108-121);
        // cell width 2 * h0 * 2 (SphSimulation.cu:46-49)
        const T cellWidth = (T(2) * p.h0) * T(2);
        const V3<T> scale {p.dmax.x - p.dmin.x, p.dmax.y - p.dmin.y, p.dmax.z - p.dmin.z};
        gx = static_cast<int>(std::ceil(scale.x / cellWidth));
        gy = static_cast<int>(std::ceil(scale.y / cellWidth));
        gz = static_cast<int>(std::ceil(scale.z / cellWidth));
        cs = {scale.x / static_cast<T>(gx), scale.y / static_cast<T>(gy), scale.z / static_cast<T>(gz)};
    }

    uint32_t cellCount() const { return static_cast<uint32_t>(gx * gy * gz); }

    // calculateCellIndex + flattenCellIndex (NeighborGrid.cuh:166-177)
    int32_t cellKey(const V4<T>& q) const
    {
        const T scx = p.dmax.x - p.dmin.x;
        const T scy = p.dmax.y - p.dmin.y;
        const T scz = p.dmax.z - p.dmin.z;
        const T rx = gmin(gmax(q.x - p.dmin.x, T(0)), scx);
        const T ry = gmin(gmax(q.y - p.dmin.y, T(0)), scy);
        const T rz = gmin(gmax(q.z - p.dmin.z, T(0)), scz);
        const int cx = std::min(static_cast<int>(rx / cs.x), gx - 1);
        const int cy = std::min(static_cast<int>(ry / cs.y), gy - 1);
        const int cz = std::min(static_cast<int>(rz / cs.z), gz - 1);
        return cx + (cy * gx) + (cz * gx * gy);
    }

    // NeighborGrid::updateFluid / updateBoundary (NeighborGrid.cu:33-49):
    // reset, assign, stable sort_by_key, calculateCellStartAndEndIndices
    void buildGrid(Grid& g, const std::vector<V4<T>>& points, uint32_t count) const
    {
        g.count = count;
        const uint32_t cells = cellCount();
        g.cellStart.assign(cells, -1);
        g.cellEnd.assign(cells, -1);
        std::vector<int32_t> keys(count);
        for (uint32_t i = 0; i < count; ++i)
        {
            keys[i] = cellKey(points[i]);
        }
        std::vector<int32_t> order(count);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return keys[a] < keys[b]; });
        g.keys.resize(count);
        g.idx.resize(count);
        for (uint32_t s = 0; s < count; ++s)
        {
            g.idx[s] = order[s];
            g.keys[s] = keys[order[s]];
        }
        // calculateCellStartAndEndIndices (NeighborKernels.cu:17-44), one
        // iteration per thread; the writes of different threads never collide.
        const bool fixQuirk = (p.flags & ASPH_FLAG_FIX_CELL_START) != 0;
        for (uint32_t s = 0; s < count; ++s)
        {
            const int32_t c = g.keys[s];
            if (s == count - 1)
            {
                g.cellEnd[c] = static_cast<int32_t>(s);
                if (fixQuirk && s == 0)
                {
                    g.cellStart[c] = 0;
                }
                if (fixQuirk && s > 0 && g.keys[s - 1] != c)
                {
                    g.cellStart[c] = static_cast<int32_t>(s);
                }
                continue;
            }
            if (s == 0)
            {
                g.cellStart[c] = 0;
                if (!fixQuirk)
                {
                    continue;
                }
            }
            const int32_t cn = g.keys[s + 1];
            if (c != cn)
            {
                g.cellStart[cn] = static_cast<int32_t>(s + 1);
                g.cellEnd[c] = static_cast<int32_t>(s);
            }
        }
    }

    // forEachFluidNeighbor / forEachBoundaryNeighbor (NeighborGrid.cuh:90-164)
    template <class F>
    void forEach(const Grid& g, const std::vector<V4<T>>& points, const V4<T>& q, T R, F&& f) const
    {
        const int x0 = std::max(static_cast<int>(((q.x - p.dmin.x) - R) / cs.x), 0);
        const int y0 = std::max(static_cast<int>(((q.y - p.dmin.y) - R) / cs.y), 0);
        const int z0 = std::max(static_cast<int>(((q.z - p.dmin.z) - R) / cs.z), 0);
        const int x1 = std::min(static_cast<int>(((q.x - p.dmin.x) + R) / cs.x), gx - 1);
        const int y1 = std::min(static_cast<int>(((q.y - p.dmin.y) + R) / cs.y), gy - 1);
        const int z1 = std::min(static_cast<int>(((q.z - p.dmin.z) + R) / cs.z), gz - 1);
        for (int x = x0; x <= x1; ++x)
        {
            for (int y = y0; y <= y1; ++y)
            {
                for (int z = z0; z <= z1; ++z)
                {
                    const int c = x + (y * gx) + (z * gx * gy);
                    const int s = g.cellStart[c];
                    const int e = g.cellEnd[c];
                    if (s == -1 || s > e)
                    {
                        continue;
                    }
                    for (int i = s; i <= e; ++i)
                    {
                        const int j = g.idx[i];
                        f(static_cast<uint32_t>(j), points[j]);
                    }
                }
            }
        }
    }

    // -------------------------------------------------------- physics --
    V3<T> gravityAt(double t) const
    {
        V3<T> g = p.g;
        if (p.sloshA != T(0) && p.sloshT != T(0))
        {
            double s = 0.0;
            double c = 0.0;
            det_sincos(6.283185307179586 * t / static_cast<double>(p.sloshT), &s, &c);
            g.x = static_cast<T>(static_cast<double>(p.g.x) - static_cast<double>(p.sloshA) * s);
        }
        return g;
    }

    // halfKickVelocities (Algorithm.cu:270-284)
    void halfKick(T halfDt)
    {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < static_cast<int64_t>(n); ++i)
        {
            V4<T> v = vel[i];
            const V4<T>& a = acc[i];
            v = {v.x + a.x * halfDt, v.y + a.y * halfDt, v.z + a.z * halfDt, v.w + a.w * halfDt};
            const T len = std::sqrt(dot4(v, v));
            if (len > p.vmax)
            {
                const T s = p.vmax / len;
                v = {v.x * s, v.y * s, v.z * s, v.w * s};
            }
            vel[i] = v;
        }
    }

    // updatePositions (Algorithm.cu:286-298)
    void drift(T dt)
    {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < static_cast<int64_t>(n); ++i)
        {
            V4<T> q = pos[i];
            const V4<T>& v = vel[i];
            q = {q.x + v.x * dt, q.y + v.y * dt, q.z + v.z * dt, q.w + v.w * dt};
            q.x = gmin(gmax(q.x, p.dmin.x), p.dmax.x);
            q.y = gmin(gmax(q.y, p.dmin.y), p.dmax.y);
            q.z = gmin(gmax(q.z, p.dmin.z), p.dmax.z);
            pos[i] = q;
        }
    }

    // computeDensities (Algorithm.cu:17-54) + computeBoundaryDensityContribution (BoundaryAlgorithm.cu:15-51)
    void densities()
    {
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t ii = 0; ii < static_cast<int64_t>(n); ++ii)
        {
            const auto i = static_cast<uint32_t>(ii);
            const V4<T> q = pos[i];
            const T hi = h[i];
            T density = T(0);
            forEach(fluid, pos, q, T(2) * hi, [&](uint32_t j, const V4<T>& pj) {
                const V4<T> off = sub4(pj, q);
                const T d2 = dot4(off, off);
                const T hj = h[j];
                const T r2 = (T(2) * hj) * (T(2) * hj);
                if (d2 > r2)
                {
                    return;
                }
                const T d = std::sqrt(d2);
                density = density + m[j] * wendland(d, hj);
            });
            rho[i] = density;
        }
        if (nb == 0)
        {
            return;
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t ii = 0; ii < static_cast<int64_t>(n); ++ii)
        {
            const auto i = static_cast<uint32_t>(ii);
            const V4<T> q = pos[i];
            const T hi = h[i];
            const T r2 = (T(2) * hi) * (T(2) * hi);
            T density = T(0);
            forEach(boundary, bpos, q, T(2) * hi, [&](uint32_t b, const V4<T>& pb) {
                const V4<T> off = sub4(pb, q);
                const T d2 = dot4(off, off);
                if (d2 > r2)
                {
                    return;
                }
                const T d = std::sqrt(d2);
                density = density + bpsi[b] * wendland(d, hi);
            });
            rho[i] = rho[i] + density;
        }
    }

    // computeExternalAccelerations, computePressureAccelerations,
    // computeViscosityAccelerations, computeSurfaceTensionAccelerations
    // (Algorithm.cu:56-232), computeBoundaryPressureAcceleration and
    // computeBoundaryFrictionAcceleration (BoundaryAlgorithm.cu:53-162), each
    // accumulated from zero and added to acc in the reference's launch order
    // (SphSimulation.cu:110-115).
    void forces(const V3<T>& g)
    {
        const T eps = T(0.01);
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t ii = 0; ii < static_cast<int64_t>(n); ++ii)
        {
            const auto i = static_cast<uint32_t>(ii);
            const V4<T> q = pos[i];
            const V4<T> vi = vel[i];
            const T rhoi = rho[i];
            const T hi = h[i];
            const T pi_ = taitPressure(rhoi, p.rho0, p.c0, p.gamma);
            V4<T> a {g.x, g.y, g.z, T(0)};

            V4<T> ap {};
            forEach(fluid, pos, q, T(2) * hi, [&](uint32_t j, const V4<T>& pj) {
                if (j == i)
                {
                    return;
                }
                const V4<T> off = sub4(q, pj);
                const T d2 = dot4(off, off);
                const T hj = h[j];
                const T r2 = (T(2) * hj) * (T(2) * hj);
                if (d2 > r2)
                {
                    return;
                }
                const T rhoj = rho[j];
                const T d = std::sqrt(d2);
                const V4<T> dir = d > T(0) ? V4<T> {off.x / d, off.y / d, off.z / d, off.w / d}
                                           : V4<T> {T(0), T(1), T(0), T(0)};
                const T pj_ = taitPressure(rhoj, p.rho0, p.c0, p.gamma);
                const T mj = m[j];
                const T pterm = pi_ / (rhoi * rhoi) + (pj_ / (rhoj * rhoj));
                const T dw = wendlandDerivative(d, hj);
                ap.x = ap.x - ((mj * dir.x) * dw) * pterm;
                ap.y = ap.y - ((mj * dir.y) * dw) * pterm;
                ap.z = ap.z - ((mj * dir.z) * dw) * pterm;
            });
            a = {a.x + ap.x, a.y + ap.y, a.z + ap.z, a.w};

            if (p.alpha != T(0))
            {
                V4<T> av {};
                forEach(fluid, pos, q, T(2) * hi, [&](uint32_t j, const V4<T>& pj) {
                    if (j == i)
                    {
                        return;
                    }
                    const V4<T> off = sub4(q, pj);
                    const T d2 = dot4(off, off);
                    const T hj = h[j];
                    const T r2 = (T(2) * hj) * (T(2) * hj);
                    if (d2 > r2)
                    {
                        return;
                    }
                    const V4<T> vd = sub4(vi, vel[j]);
                    const T comp = dot4(vd, off);
                    if (comp >= T(0))
                    {
                        return;
                    }
                    const T d = std::sqrt(d2);
                    const T mj = m[j];
                    const T rhoj = rho[j];
                    const T nu = (((T(2) * p.alpha) * hi) * p.c0) / (rhoi + rhoj);
                    const T piv = ((-nu) * comp) / (d2 + ((eps * hi) * hi));
                    const V4<T> dir = d > T(0) ? V4<T> {off.x / d, off.y / d, off.z / d, off.w / d}
                                               : V4<T> {T(0), T(1), T(0), T(0)};
                    const T s = (mj * piv) * wendlandDerivative(d, hj);
                    av.x = av.x - s * dir.x;
                    av.y = av.y - s * dir.y;
                    av.z = av.z - s * dir.z;
                });
                a = {a.x + av.x, a.y + av.y, a.z + av.z, a.w};
            }

            // Extension (no reference counterpart): Morris, Fox & Zhu (1997),
            // J. Comput. Phys. 136, 214, eq. 6, with mu = rho nu:
            // a_i += sum_j m_j (mu_i + mu_j) / (rho_i rho_j)
            //        * (r_ij . grad_i W_ij) / (r_ij^2 + 0.01 h_i^2) * v_ij,
            // r_ij . grad_i W_ij = |r_ij| dW/dr (W with h_j, as the reference's
            // other pair terms), added after the artificial viscosity.
            if (p.nuLam != T(0))
            {
                V4<T> am {};
                forEach(fluid, pos, q, T(2) * hi, [&](uint32_t j, const V4<T>& pj) {
                    if (j == i)
                    {
                        return;
                    }
                    const V4<T> off = sub4(q, pj);
                    const T d2 = dot4(off, off);
                    const T hj = h[j];
                    const T r2 = (T(2) * hj) * (T(2) * hj);
                    if (d2 > r2)
                    {
                        return;
                    }
                    const T d = std::sqrt(d2);
                    const T mj = m[j];
                    const T rhoj = rho[j];
                    const T k = ((mj * p.nuLam) * (rhoi + rhoj)) / (rhoi * rhoj);
                    const T s = (k * (d * wendlandDerivative(d, hj))) / (d2 + ((eps * hi) * hi));
                    const V4<T> vd = sub4(vi, vel[j]);
                    am.x = am.x + s * vd.x;
                    am.y = am.y + s * vd.y;
                    am.z = am.z + s * vd.z;
                });
                a = {a.x + am.x, a.y + am.y, a.z + am.z, a.w};
            }

            if (p.sigma != T(0))
            {
                V4<T> ast {};
                forEach(fluid, pos, q, T(2) * hi, [&](uint32_t j, const V4<T>& pj) {
                    if (j == i)
                    {
                        return;
                    }
                    const V4<T> off = sub4(q, pj);
                    const T d2 = dot4(off, off);
                    const T hj = h[j];
                    const T r2 = (T(2) * hj) * (T(2) * hj);
                    if (d2 > r2)
                    {
                        return;
                    }
                    const T d = std::sqrt(d2);
                    const T mj = m[j];
                    const V4<T> dir = d > T(0) ? V4<T> {off.x / d, off.y / d, off.z / d, off.w / d}
                                               : V4<T> {T(0), T(1), T(0), T(0)};
                    const T s = mj * wendland(d, hj);
                    ast.x = ast.x + s * dir.x;
                    ast.y = ast.y + s * dir.y;
                    ast.z = ast.z + s * dir.z;
                });
                const T k = p.sigma / m[i];
                a = {a.x + k * ast.x, a.y + k * ast.y, a.z + k * ast.z, a.w};
            }

            // boundary pressure then friction (launched unconditionally)
            {
                const T r2 = (T(2) * hi) * (T(2) * hi);
                V4<T> abp {};
                V4<T> abf {};
                if (nb > 0)
                {
                    forEach(boundary, bpos, q, T(2) * hi, [&](uint32_t b, const V4<T>& pb) {
                        const V4<T> off = sub4(q, pb);
                        const T d2 = dot4(off, off);
                        if (d2 > r2)
                        {
                            return;
                        }
                        const T d = std::sqrt(d2);
                        const V4<T> dir = d > T(0) ? V4<T> {off.x / d, off.y / d, off.z / d, off.w / d}
                                                   : V4<T> {T(0), T(1), T(0), T(0)};
                        const T psi = bpsi[b];
                        const T pterm = pi_ / (rhoi * rhoi);
                        const T s = (psi * pterm) * wendlandDerivative(d, hi);
                        abp.x = abp.x - s * dir.x;
                        abp.y = abp.y - s * dir.y;
                        abp.z = abp.z - s * dir.z;
                    });
                    forEach(boundary, bpos, q, T(2) * hi, [&](uint32_t b, const V4<T>& pb) {
                        const V4<T> off = sub4(q, pb);
                        const T d2 = dot4(off, off);
                        if (d2 > r2)
                        {
                            return;
                        }
                        const T d = std::sqrt(d2);
                        const T psi = bpsi[b];
                        const T f = bvisc[b];
                        const T comp = dot4(vi, off);
                        if (comp >= T(0))
                        {
                            return;
                        }
                        const T nu = ((f * hi) * p.c0) / (T(2) * rhoi);
                        const T piv = ((-nu) * comp) / (d2 + ((eps * hi) * hi));
                        const V4<T> dir = d > T(0) ? V4<T> {off.x / d, off.y / d, off.z / d, off.w / d}
                                                   : V4<T> {T(0), T(1), T(0), T(0)};
                        const T dw = wendlandDerivative(d, hi);
                        abf.x = abf.x - ((dir.x * psi) * piv) * dw;
                        abf.y = abf.y - ((dir.y * psi) * piv) * dw;
                        abf.z = abf.z - ((dir.z * psi) * piv) * dw;
                    });
                }
                a = {a.x + abp.x, a.y + abp.y, a.z + abp.z, a.w};
                a = {a.x + abf.x, a.y + abf.y, a.z + abf.z, a.w};
            }
            acc[i] = a;
        }
    }

    // Absolute-sum scale of each particle's fluid pair forces at the current
    // state: S_i = sum_j (|pressure term_ij| + |viscosity term_ij|), the
    // normalisation of force tolerances (interior net forces nearly cancel).
    void forceScale(double* out) const
    {
        const T eps = T(0.01);
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t ii = 0; ii < static_cast<int64_t>(n); ++ii)
        {
            const auto i = static_cast<uint32_t>(ii);
            const V4<T> q = pos[i];
            const V4<T> vi = vel[i];
            const T rhoi = rho[i];
            const T hi = h[i];
            const T pi_ = taitPressure(rhoi, p.rho0, p.c0, p.gamma);
            double S = 0.0;
            forEach(fluid, pos, q, T(2) * hi, [&](uint32_t j, const V4<T>& pj) {
                if (j == i)
                {
                    return;
                }
                const V4<T> off = sub4(q, pj);
                const T d2 = dot4(off, off);
                const T hj = h[j];
                if (d2 > (T(2) * hj) * (T(2) * hj))
                {
                    return;
                }
                const T d = std::sqrt(d2);
                const T rhoj = rho[j];
                const T dw = wendlandDerivative(d, hj);
                const T pterm = pi_ / (rhoi * rhoi) + (taitPressure(rhoj, p.rho0, p.c0, p.gamma) / (rhoj * rhoj));
                S += std::fabs(static_cast<double>(m[j] * dw * pterm));
                const V4<T> vd = sub4(vi, vel[j]);
                const T comp = dot4(vd, off);
                if (p.alpha != T(0) && comp < T(0))
                {
                    const T nu = (((T(2) * p.alpha) * hi) * p.c0) / (rhoi + rhoj);
                    const T piv = ((-nu) * comp) / (d2 + ((eps * hi) * hi));
                    S += std::fabs(static_cast<double>(m[j] * piv * dw));
                }
            });
            out[i] = S;
        }
    }

    // ------------------------------------------------------- refinement --
    bool inDomain(const V4<T>& q, T hh, T eps) const
    {
        // checkIfParticleCanBeSplitInsideDomain (criteria/Common.cuh:6-20)
        const T mo = eps * hh;
        return !(q.x - mo < p.dmin.x || q.x + mo > p.dmax.x || q.y - mo < p.dmin.y || q.y + mo > p.dmax.y ||
                 q.z - mo < p.dmin.z || q.z + mo > p.dmax.z);
    }

    V3<T> minSurfaceDistance(const V4<T>& q) const
    {
        // calculateMinSurfaceDistance (InterfaceCriterion.cuh:52-67)
        const T mx = gmin(q.x - p.dmin.x, p.dmax.x - q.x);
        const T my = gmin(q.y - p.dmin.y, p.dmax.y - q.y);
        const T mz = gmin(q.z - p.dmin.z, p.dmax.z - q.z);
        return {gmax(T(0), mx), gmax(T(0), my), gmax(T(0), mz)};
    }

    V3<T> zoneCenterAt(double t) const
    {
        V3<T> c = ref.zoneCenter;
        if (ref.zonePathPeriod != T(0))
        {
            double s = 0.0;
            double co = 0.0;
            det_sincos(6.283185307179586 * t / static_cast<double>(ref.zonePathPeriod), &s, &co);
            c.x = static_cast<T>(static_cast<double>(ref.zoneCenter.x) + static_cast<double>(ref.zonePathRadius) * co);
            c.z = static_cast<T>(static_cast<double>(ref.zoneCenter.z) + static_cast<double>(ref.zonePathRadius) * s);
        }
        return c;
    }

    // vorticity estimate shared by both vorticity generators (VorticityCriterion.cuh:67-113)
    bool vorticity(const Grid& g, uint32_t id, T* mag) const
    {
        const V4<T> q = pos[id];
        const T hh = h[id];
        T dxx = 0, dxy = 0, dxz = 0, dyx = 0, dyy = 0, dyz = 0, dzx = 0, dzy = 0, dzz = 0;
        int valid = 0;
        forEach(g, pos, q, hh, [&](uint32_t j, const V4<T>& pj) {
            if (j == id)
            {
                return;
            }
            const V3<T> r {pj.x - q.x, pj.y - q.y, pj.z - q.z};
            const T dist = std::sqrt(dot3(r, r));
            if (dist < hh && dist > T(1e-6F))
            {
                const V4<T> vj = vel[j];
                const V4<T> vid = vel[id];
                const V3<T> vd {vj.x - vid.x, vj.y - vid.y, vj.z - vid.z};
                const T gw = wendlandDerivative(dist, hh);
                const T w = m[j] / rho[j];
                const V3<T> gd {r.x / dist, r.y / dist, r.z / dist};
                dxx = dxx + ((w * vd.x) * gd.x) * gw;
                dxy = dxy + ((w * vd.x) * gd.y) * gw;
                dxz = dxz + ((w * vd.x) * gd.z) * gw;
                dyx = dyx + ((w * vd.y) * gd.x) * gw;
                dyy = dyy + ((w * vd.y) * gd.y) * gw;
                dyz = dyz + ((w * vd.y) * gd.z) * gw;
                dzx = dzx + ((w * vd.z) * gd.x) * gw;
                dzy = dzy + ((w * vd.z) * gd.y) * gw;
                dzz = dzz + ((w * vd.z) * gd.z) * gw;
                valid++;
            }
        });
        if (valid < 4)
        {
            return false;
        }
        const V3<T> om {dzy - dyz, dxz - dzx, dyx - dxy};
        *mag = std::sqrt(dot3(om, om));
        return true;
    }

    // Split generators (VelocityCriterion.cuh:54-78, InterfaceCriterion.cuh:69-95,
    // VorticityCriterion.cuh:53-125) + the zone extension.
    T splitCriterion(uint32_t id, const Grid& g, const V3<T>& zc) const
    {
        const T minMass = ref.minMassRatio * p.m0;
        if (m[id] < minMass)
        {
            return T(-1);
        }
        const V4<T> q = pos[id];
        const T hh = h[id];
        switch (ref.criterion)
        {
        case ASPH_CRITERION_INTERFACE: {
            const V3<T> md = minSurfaceDistance(q);
            const T mo = ref.eps * hh;
            if (gmin(gmin(md.x, md.y), md.z) < mo)
            {
                return T(-1);
            }
            const V3<T> thr {ref.ifSplit * (p.dmax.x - p.dmin.x), ref.ifSplit * (p.dmax.y - p.dmin.y),
                             ref.ifSplit * (p.dmax.z - p.dmin.z)};
            const V3<T> nd {md.x / thr.x, md.y / thr.y, md.z / thr.z};
            return T(1) - gmin(gmin(nd.x, nd.y), nd.z);
        }
        case ASPH_CRITERION_VORTICITY: {
            if (!inDomain(q, hh, ref.eps))
            {
                return T(-1);
            }
            T mag = 0;
            if (!vorticity(g, id, &mag))
            {
                return T(-1);
            }
            return mag > ref.vortSplit ? mag : T(-1);
        }
        case ASPH_CRITERION_ZONE: {
            if (!inDomain(q, hh, ref.eps))
            {
                return T(-1);
            }
            const V3<T> r {q.x - zc.x, q.y - zc.y, q.z - zc.z};
            const T d = std::sqrt(dot3(r, r));
            return d < ref.zoneRadius ? T(2) - d / ref.zoneRadius : T(-1);
        }
        default: {
            if (!inDomain(q, hh, ref.eps))
            {
                return T(-1);
            }
            const V4<T> v = vel[id];
            const V3<T> v3 {v.x, v.y, v.z};
            const T mag = std::sqrt(dot3(v3, v3));
            if (mag < ref.velSplit)
            {
                return T(-1);
            }
            return mag;
        }
        }
    }

    // Merge generators (VelocityCriterion.cuh:80-101, InterfaceCriterion.cuh:97-115,
    // VorticityCriterion.cuh:127-195) + the zone extension.
    T mergeCriterion(uint32_t id, const Grid& g, const V3<T>& zc) const
    {
        const T maxMass = ref.maxMassRatio * p.m0;
        if (m[id] > maxMass)
        {
            return T(-1);
        }
        const V4<T> q = pos[id];
        switch (ref.criterion)
        {
        case ASPH_CRITERION_INTERFACE: {
            const V3<T> md = minSurfaceDistance(q);
            const V3<T> thr {ref.ifMerge * (p.dmax.x - p.dmin.x), ref.ifMerge * (p.dmax.y - p.dmin.y),
                             ref.ifMerge * (p.dmax.z - p.dmin.z)};
            const V3<T> nd {md.x / thr.x, md.y / thr.y, md.z / thr.z};
            return gmin(gmin(nd.x, nd.y), nd.z) - T(1);
        }
        case ASPH_CRITERION_VORTICITY: {
            T mag = 0;
            if (!vorticity(g, id, &mag))
            {
                return T(-1);
            }
            return mag < ref.vortMerge ? mag / ref.vortMerge : T(-1);
        }
        case ASPH_CRITERION_ZONE: {
            const V3<T> r {q.x - zc.x, q.y - zc.y, q.z - zc.z};
            const T d = std::sqrt(dot3(r, r));
            return d > ref.zoneRadius ? ref.zoneRadius / d : T(2);
        }
        default: {
            const V4<T> v = vel[id];
            const V3<T> v3 {v.x, v.y, v.z};
            const T mag = std::sqrt(dot3(v3, v3));
            if (mag < ref.velMerge)
            {
                return mag / ref.velMerge;
            }
            return T(-1);
        }
        }
    }

    // findTopParticlesToMerge / findTopParticlesToSplit (AdaptiveAlgorithm.cuh:56-131):
    // ids sorted stably by value (ascending or descending), count of values > 0,
    // k = min(count, round(N * ratio)).
    std::vector<uint32_t> topList(const std::vector<T>& values, bool descending, uint32_t* kOut) const
    {
        const uint32_t count = static_cast<uint32_t>(values.size());
        std::vector<uint32_t> ids(count);
        std::iota(ids.begin(), ids.end(), 0U);
        if (descending)
        {
            std::stable_sort(ids.begin(), ids.end(), [&](uint32_t a, uint32_t b) { return values[a] > values[b]; });
        }
        else
        {
            std::stable_sort(ids.begin(), ids.end(), [&](uint32_t a, uint32_t b) { return values[a] < values[b]; });
        }
        uint32_t positive = 0;
        for (const T v : values)
        {
            positive += v > T(0) ? 1U : 0U;
        }
        const float limit = std::round(static_cast<float>(count) * static_cast<float>(ref.batchRatio));
        const uint32_t k = std::min(positive, static_cast<uint32_t>(limit));
        ids.resize(k);
        *kOut = k;
        return ids;
    }

    // calculateMergedSmoothingLength (AdaptiveAlgorithm.cu:50-67)
    T mergedSmoothingLength(const V4<T>& pa, const V4<T>& pb, T ma, T mb, T ha, T hb, const V4<T>& pm, T M) const
    {
        const V3<T> ra {pa.x - pm.x, pa.y - pm.y, pa.z - pm.z};
        const V3<T> rb {pb.x - pm.x, pb.y - pm.y, pb.z - pm.z};
        const T da = std::sqrt(dot3(ra, ra));
        const T db = std::sqrt(dot3(rb, rb));
        const T wa = wendland(da, ha);
        const T wb = wendland(db, hb);
        const T denom = (ma * wa) + (mb * wb);
        // ASPH_FLAG_KERNEL_MERGED_H: the Wendland normalisation (WendlandKernel.cuh:11)
        const T coef = (p.flags & ASPH_FLAG_KERNEL_MERGED_H) != 0U ? Const<T>::wcoef : Const<T>::mergeCoef;
        return Const<T>::cbrt((coef * M) / denom);
    }

    void performAdaptiveRefinement(double t)
    {
        trace = Trace {};
        trace.t.refined = 1;
        const uint32_t before = n;
        trace.t.count_before = before;
        const V3<T> zc = zoneCenterAt(t);
        const uint32_t noCandidate = 0xFFFFFFFFU;

        // ---- identifyAndMergeParticles (AdaptiveSphSimulation.cu:195-252)
        if (n > 1)
        {
            std::vector<T> values(n);
#pragma omp parallel for schedule(dynamic, 256)
            for (int64_t i = 0; i < static_cast<int64_t>(n); ++i)
            {
                values[i] = mergeCriterion(static_cast<uint32_t>(i), fluid, zc);
            }
            uint32_t k = 0;
            const std::vector<uint32_t> list = topList(values, false, &k);
            trace.mergeList = list;
            trace.t.merge_eligible = k;
            std::vector<uint32_t> cand(n, 0U);  // reset value 0 (AdaptiveSphSimulation.cu:127)
            if (k > 0)
            {
                const T massLimit = ref.maxMassRatio * p.m0;
                // identifyMergeCandidates (AdaptiveAlgorithm.cu:159-206)
#pragma omp parallel for schedule(dynamic, 64)
                for (int64_t e = 0; e < static_cast<int64_t>(k); ++e)
                {
                    const uint32_t id = list[e];
                    const V4<T> q = pos[id];
                    const T hh = h[id];
                    T closest = static_cast<T>(FLT_MAX);
                    uint32_t best = noCandidate;
                    forEach(fluid, pos, q, hh, [&](uint32_t j, const V4<T>& pj) {
                        if (j == id)
                        {
                            return;
                        }
                        const V4<T> off = sub4(pj, q);
                        const T d2 = dot4(off, off);
                        const T hmin = gmin(hh, h[j]);
                        if (d2 < hmin * hmin && d2 < closest && m[j] < massLimit)
                        {
                            closest = d2;
                            best = j;
                        }
                    });
                    cand[id] = best;
                }
                // resolveMergePairs (AdaptiveAlgorithm.cu:208-234)
                std::vector<uint32_t> removal(n, 0U);
                std::vector<uint32_t> keeps;
                for (uint32_t i = 0; i < n; ++i)
                {
                    const uint32_t c = cand[i];
                    if (c == noCandidate || removal[i] != 0U)
                    {
                        continue;
                    }
                    if (cand[c] == i && i < c)
                    {
                        keeps.push_back(i);
                        removal[c] = 1U;
                    }
                }
                trace.candidates = cand;
                trace.pairKeep = keeps;
                trace.t.merge_pairs = static_cast<uint32_t>(keeps.size());
                if (!keeps.empty())
                {
                    // performMerges (AdaptiveAlgorithm.cu:236-275)
                    for (const uint32_t a : keeps)
                    {
                        const uint32_t b = cand[a];
                        const V4<T> pa = pos[a], pb = pos[b];
                        const T ma = m[a], mb = m[b];
                        const T ha = h[a], hb = h[b];
                        const V4<T> va = vel[a], vb = vel[b];
                        const V4<T> aa = acc[a], ab = acc[b];
                        const T M = ma + mb;
                        const V4<T> np {(ma * pa.x + mb * pb.x) / M, (ma * pa.y + mb * pb.y) / M,
                                        (ma * pa.z + mb * pb.z) / M, (ma * pa.w + mb * pb.w) / M};
                        const V4<T> nv {(ma * va.x + mb * vb.x) / M, (ma * va.y + mb * vb.y) / M,
                                        (ma * va.z + mb * vb.z) / M, (ma * va.w + mb * vb.w) / M};
                        const V4<T> na {(ma * aa.x + mb * ab.x) / M, (ma * aa.y + mb * ab.y) / M,
                                        (ma * aa.z + mb * ab.z) / M, (ma * aa.w + mb * ab.w) / M};
                        pos[a] = np;
                        vel[a] = nv;
                        acc[a] = na;
                        m[a] = M;
                        rad[a] = p.r0 * Const<T>::cbrt(M / p.m0);
                        h[a] = mergedSmoothingLength(pa, pb, ma, mb, ha, hb, np, M);
                        rho[a] = T(0);
                    }
                    // exclusive_scan + compactParticles (AdaptiveSphSimulation.cu:240-249,
                    // AdaptiveAlgorithm.cu:119-157): stable compaction
                    uint32_t w = 0;
                    for (uint32_t i = 0; i < n; ++i)
                    {
                        if (removal[i] == 0U)
                        {
                            pos[w] = pos[i];
                            vel[w] = vel[i];
                            acc[w] = acc[i];
                            m[w] = m[i];
                            rad[w] = rad[i];
                            h[w] = h[i];
                            rho[w] = rho[i];
                            ++w;
                        }
                    }
                    n = w;
                }
            }
            else
            {
                trace.candidates = cand;
            }
        }
        trace.t.count_after_merge = n;

        // ---- identifyAndSplitParticles (AdaptiveSphSimulation.cu:135-193)
        const uint32_t removed = before - n;
        if (removed > 0)
        {
            std::vector<T> values(n);
#pragma omp parallel for schedule(dynamic, 256)
            for (int64_t i = 0; i < static_cast<int64_t>(n); ++i)
            {
                // vorticity walks the grid of this step (built before compaction)
                values[i] = splitCriterion(static_cast<uint32_t>(i), fluid, zc);
            }
            uint32_t k = 0;
            const std::vector<uint32_t> list = topList(values, true, &k);
            trace.t.split_selected = k;
            const uint32_t toSplit = std::min(k, removed / 12U);
            if (toSplit > 0)
            {
                const uint32_t tpb = p.tpb;
                const uint32_t launched = ((toSplit + tpb - 1U) / tpb) * tpb;
                const uint32_t effective = std::min(k, launched);
                // splitParticles (AdaptiveAlgorithm.cu:69-117); slots assigned in
                // list order (the reference's atomicAdd order is unspecified)
                for (uint32_t r = 0; r < effective; ++r)
                {
                    if (static_cast<uint64_t>(n) + 12U > ref.maxCount)
                    {
                        break;
                    }
                    const uint32_t parent = list[r];
                    const uint32_t base = n;
                    n += 12U;
                    splitOne(parent, base);
                    trace.splitParents.push_back(parent);
                }
            }
        }
        trace.t.split_performed = static_cast<uint32_t>(trace.splitParents.size());
        trace.t.count_after = n;
    }

    void splitOne(uint32_t parent, uint32_t base)
    {
        // icosahedronVertices (AdaptiveAlgorithm.cu:24-48): phi * invnorm in FP32
        constexpr float phi = 1.618033988749894848204586834365638F;
        constexpr float inv = 0.5257311121F;
        constexpr float pin = phi * inv;
        const float verts[12][3] = {{0.F, pin, inv},  {0.F, pin, -inv},  {0.F, -pin, -inv}, {0.F, -pin, inv},
                                    {inv, 0.F, pin},  {-inv, 0.F, pin},  {inv, 0.F, -pin},  {-inv, 0.F, -pin},
                                    {pin, inv, 0.F},  {pin, -inv, 0.F},  {-pin, inv, 0.F},  {-pin, -inv, 0.F}};
        const V4<T> op = pos[parent];
        const V4<T> ov = vel[parent];
        const V4<T> oa = acc[parent];
        const T om = m[parent];
        const T orad = rad[parent];
        const T oh = h[parent];
        const T dm = om * ref.vertexRatio;
        const T cm = om * ref.centerRatio;
        const T newRadius = orad * Const<T>::cbrt(dm / om);
        const T newH = ref.alphaSplit * oh;
        m[parent] = cm;
        rad[parent] = orad * Const<T>::cbrt(cm / om);
        h[parent] = newH;
        for (uint32_t v = 0; v < 12; ++v)
        {
            const uint32_t ni = base + v;
            const T ox = (static_cast<T>(verts[v][0]) * ref.eps) * oh;
            const T oy = (static_cast<T>(verts[v][1]) * ref.eps) * oh;
            const T oz = (static_cast<T>(verts[v][2]) * ref.eps) * oh;
            pos[ni] = {op.x + ox, op.y + oy, op.z + oz, op.w + T(0)};
            vel[ni] = ov;
            acc[ni] = oa;
            m[ni] = dm;
            rad[ni] = newRadius;
            h[ni] = newH;
            rho[ni] = T(0);
        }
    }

    // SphSimulation::update (SphSimulation.cu:101-120) and
    // AdaptiveSphSimulation::update (AdaptiveSphSimulation.cu:59-85)
    void update(T dt)
    {
        const T halfDt = dt / T(2);
        trace.t.refined = 0;
        halfKick(halfDt);
        drift(dt);
        const double tEval = time + static_cast<double>(dt);
        buildGrid(fluid, pos, n);
        densities();
        forces(gravityAt(tEval));
        if (ref.enabled)
        {
            const uint32_t ic = ref.initialCooldown;
            if (frame == ic || (frame > ic && frame % ref.cooldown == 0))
            {
                performAdaptiveRefinement(tEval);
            }
        }
        halfKick(halfDt);
        if (ref.enabled)
        {
            frame++;
        }
        time = tEval;
    }

    // sumAllNeighbors (Algorithm.cu:234-268)
    void neighborSets(uint32_t* counts, uint64_t* prints) const
    {
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t ii = 0; ii < static_cast<int64_t>(n); ++ii)
        {
            const auto i = static_cast<uint32_t>(ii);
            const V4<T> q = pos[i];
            uint32_t c = 0;
            uint64_t hsh = 0;
            forEach(fluid, pos, q, T(2) * h[i], [&](uint32_t j, const V4<T>& pj) {
                if (j == i)
                {
                    return;
                }
                const V4<T> off = sub4(pj, q);
                const T d2 = dot4(off, off);
                const T r2 = (T(2) * h[j]) * (T(2) * h[j]);
                if (d2 <= r2)
                {
                    c++;
                    hsh += mix64(j);
                }
            });
            counts[i] = c;
            if (prints != nullptr)
            {
                prints[i] = hsh;
            }
        }
    }
};

// StaticBoundaryDomain::generateWallParticles (StaticBoundaryDomain.cu:88-147)
inline std::vector<V4<float>> wallParticles(const asph_domain& d, float spacing)
{
    std::vector<V4<float>> out;
    const float off = 0.05F;
    const V3<float> mn {d.min.x - off, d.min.y - off, d.min.z - off};
    const V3<float> mx {d.max.x + off, d.max.y + off, d.max.z + off};
    for (float x = mn.x; x < mx.x; x += spacing)
    {
        for (float z = mn.z; z < mx.z; z += spacing)
        {
            out.push_back({x, mn.y, z, 0.F});
        }
    }
    for (float x = mn.x; x < mx.x; x += spacing)
    {
        for (float z = mn.z; z < mx.z; z += spacing)
        {
            out.push_back({x, mx.y, z, 0.F});
        }
    }
    for (float y = mn.y + spacing; y < mx.y; y += spacing)
    {
        for (float z = mn.z; z < mx.z; z += spacing)
        {
            out.push_back({mn.x, y, z, 0.F});
        }
    }
    for (float y = mn.y + spacing; y < mx.y; y += spacing)
    {
        for (float z = mn.z; z < mx.z; z += spacing)
        {
            out.push_back({mx.x, y, z, 0.F});
        }
    }
    for (float x = mn.x + spacing; x < mx.x; x += spacing)
    {
        for (float y = mn.y + spacing; y < mx.y; y += spacing)
        {
            out.push_back({x, y, mn.z, 0.F});
        }
    }
    for (float x = mn.x + spacing; x < mx.x; x += spacing)
    {
        for (float y = mn.y + spacing; y < mx.y; y += spacing)
        {
            out.push_back({x, y, mx.z, 0.F});
        }
    }
    return out;
}

// StaticBoundaryDomain::computePsi (StaticBoundaryDomain.cu:149-174): delta =
// sum of W(sqrt(d2), h) over d2 < (2h)^2 (nanoflann's strict radius test),
// d2 accumulated as ((0 + dx^2) + dy^2) + dz^2 with dx = query - point, summed
// in ascending point index (nanoflann's traversal order is unpinned).
// Brute force over a uniform bucket grid of width 2h.
inline std::vector<float> wallPsi(const std::vector<V4<float>>& pts, float h, float rho0)
{
    const size_t count = pts.size();
    std::vector<float> psi(count);
    if (count == 0)
    {
        return psi;
    }
    const float R = 2.0F * h;
    const float R2 = R * R;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (const auto& q : pts)
    {
        lo[0] = std::min(lo[0], q.x);
        lo[1] = std::min(lo[1], q.y);
        lo[2] = std::min(lo[2], q.z);
        hi[0] = std::max(hi[0], q.x);
        hi[1] = std::max(hi[1], q.y);
        hi[2] = std::max(hi[2], q.z);
    }
    const double w = static_cast<double>(R);
    int dims[3];
    for (int a = 0; a < 3; ++a)
    {
        dims[a] = std::max(1, static_cast<int>(std::floor((static_cast<double>(hi[a]) - lo[a]) / w)) + 1);
    }
    auto bucketOf = [&](const V4<float>& q, int a) {
        const double v = (a == 0 ? q.x : (a == 1 ? q.y : q.z));
        int b = static_cast<int>(std::floor((v - lo[a]) / w));
        return std::min(std::max(b, 0), dims[a] - 1);
    };
    const size_t nbk = static_cast<size_t>(dims[0]) * dims[1] * dims[2];
    std::vector<uint32_t> start(nbk + 1, 0);
    std::vector<uint32_t> bucket(count);
    for (size_t i = 0; i < count; ++i)
    {
        const size_t b = bucketOf(pts[i], 0) + dims[0] * (bucketOf(pts[i], 1) + static_cast<size_t>(dims[1]) * bucketOf(pts[i], 2));
        bucket[i] = static_cast<uint32_t>(b);
        start[b + 1]++;
    }
    for (size_t b = 0; b < nbk; ++b)
    {
        start[b + 1] += start[b];
    }
    std::vector<uint32_t> items(count);
    {
        std::vector<uint32_t> fill(start.begin(), start.end() - 1);
        for (size_t i = 0; i < count; ++i)
        {
            items[fill[bucket[i]]++] = static_cast<uint32_t>(i);
        }
    }
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t ii = 0; ii < static_cast<int64_t>(count); ++ii)
    {
        const auto& q = pts[ii];
        const int bx = bucketOf(q, 0), by = bucketOf(q, 1), bz = bucketOf(q, 2);
        std::vector<uint32_t> hits;
        for (int z = std::max(bz - 1, 0); z <= std::min(bz + 1, dims[2] - 1); ++z)
        {
            for (int y = std::max(by - 1, 0); y <= std::min(by + 1, dims[1] - 1); ++y)
            {
                for (int x = std::max(bx - 1, 0); x <= std::min(bx + 1, dims[0] - 1); ++x)
                {
                    const size_t b = x + dims[0] * (y + static_cast<size_t>(dims[1]) * z);
                    for (uint32_t s = start[b]; s < start[b + 1]; ++s)
                    {
                        const uint32_t j = items[s];
                        const auto& o = pts[j];
                        float d2 = 0.F;
                        const float dx = q.x - o.x;
                        d2 = d2 + dx * dx;
                        const float dy = q.y - o.y;
                        d2 = d2 + dy * dy;
                        const float dz = q.z - o.z;
                        d2 = d2 + dz * dz;
                        if (d2 < R2)
                        {
                            hits.push_back(j);
                        }
                    }
                }
            }
        }
        std::sort(hits.begin(), hits.end());
        float sum = 0.F;
        for (const uint32_t j : hits)
        {
            const auto& o = pts[j];
            float d2 = 0.F;
            const float dx = q.x - o.x;
            d2 = d2 + dx * dx;
            const float dy = q.y - o.y;
            d2 = d2 + dy * dy;
            const float dz = q.z - o.z;
            d2 = d2 + dz * dz;
            sum = sum + wendland(std::sqrt(d2), h);
        }
        const float volume = 1.0F / sum;
        psi[ii] = rho0 * volume;
    }
    return psi;
}

}  // namespace oracle

// ================================================================ C API ==

namespace
{

struct OracleHandle
{
    int precision = 32;
    std::unique_ptr<oracle::Sim<float>> f;
    std::unique_ptr<oracle::Sim<double>> d;
};

template <class T>
void loadParams(oracle::Sim<T>& s, const asph_parameters& p, const asph_refinement_parameters* r)
{
    auto& q = s.p;
    q.dmin = {T(p.domain.min.x), T(p.domain.min.y), T(p.domain.min.z)};
    q.dmax = {T(p.domain.max.x), T(p.domain.max.y), T(p.domain.max.z)};
    q.friction = p.domain.friction;
    q.g = {T(p.gravity.x), T(p.gravity.y), T(p.gravity.z)};
    q.c0 = p.speed_of_sound;
    q.rho0 = p.rest_density;
    q.restitution = p.restitution;
    q.alpha = p.viscosity_constant;
    q.sigma = p.surface_tension_constant;
    q.vmax = p.max_velocity;
    q.h0 = p.base_smoothing_radius;
    q.r0 = p.base_particle_radius;
    q.m0 = p.base_particle_mass;
    q.tpb = p.threads_per_block;
    q.gamma = p.tait_gamma;
    q.sloshA = p.slosh_amplitude;
    q.sloshT = p.slosh_period;
    q.flags = p.flags;
    q.nuLam = p.laminar_viscosity;
    auto& f = s.ref;
    f.enabled = r != nullptr && r->enabled != 0U;
    if (r != nullptr)
    {
        f.minMassRatio = r->min_mass_ratio;
        f.maxMassRatio = r->max_mass_ratio;
        f.maxCount = r->max_particle_count;
        f.batchRatio = r->max_batch_ratio;
        f.initialCooldown = r->initial_cooldown;
        f.cooldown = r->cooldown;
        f.criterion = r->criterion_type;
        f.eps = r->splitting.epsilon;
        f.alphaSplit = r->splitting.alpha;
        f.centerRatio = r->splitting.center_mass_ratio;
        f.vertexRatio = r->splitting.vertex_mass_ratio;
        f.velSplit = r->velocity.split;
        f.velMerge = r->velocity.merge;
        f.vortSplit = r->vorticity.split;
        f.vortMerge = r->vorticity.merge;
        f.ifSplit = r->interface_ratio.split;
        f.ifMerge = r->interface_ratio.merge;
        f.zoneCenter = {T(r->zone.center.x), T(r->zone.center.y), T(r->zone.center.z)};
        f.zoneRadius = r->zone.radius;
        f.zonePathRadius = r->zone.path_radius;
        f.zonePathPeriod = r->zone.path_period;
    }
}

template <class T>
std::unique_ptr<oracle::Sim<T>> makeSim(const asph_parameters& p, const asph_vec4* positions, uint32_t count,
                                        const asph_boundary_particle* boundary, uint32_t nb,
                                        const asph_refinement_parameters* r, uint32_t maxFluid, uint32_t maxBoundary)
{
    auto s = std::make_unique<oracle::Sim<T>>();
    loadParams(*s, p, r);
    // SphSimulation ctor (SphSimulation.cu:29-80); the adaptive variant uses
    // maxParticleCount as capacity (AdaptiveSphSimulation.cu:37)
    const uint32_t cap = s->ref.enabled ? s->ref.maxCount : maxFluid;
    if (count > cap)
    {
        throw std::runtime_error("particle count exceeds capacity");
    }
    if (nb > maxBoundary)
    {
        throw std::runtime_error("boundary count exceeds capacity");
    }
    s->n = count;
    s->cap = cap;
    s->bcap = maxBoundary;
    s->pos.assign(cap, {});
    s->vel.assign(cap, {});
    s->acc.assign(cap, {s->p.g.x, s->p.g.y, s->p.g.z, T(0)});
    s->rho.assign(cap, T(0));
    s->rad.assign(cap, T(0));
    s->h.assign(cap, T(0));
    s->m.assign(cap, T(0));
    for (uint32_t i = 0; i < count; ++i)
    {
        s->pos[i] = {T(positions[i].x), T(positions[i].y), T(positions[i].z), T(positions[i].w)};
        s->rad[i] = s->p.r0;
        s->h[i] = s->p.h0;
        s->m[i] = s->p.m0;
    }
    s->nb = nb;
    s->bpos.resize(nb);
    s->bpsi.resize(nb);
    s->bvisc.resize(nb);
    s->brad.resize(nb);
    for (uint32_t i = 0; i < nb; ++i)
    {
        const auto& b = boundary[i];
        s->bpos[i] = {T(b.position.x), T(b.position.y), T(b.position.z), T(b.position.w)};
        s->bpsi[i] = b.psi;
        s->bvisc[i] = s->p.friction;
        s->brad[i] = s->p.r0;
    }
    s->createGrid();
    s->buildGrid(s->boundary, s->bpos, nb);
    s->buildGrid(s->fluid, s->pos, 0);
    return s;
}

thread_local std::string g_err;

}  // namespace

extern "C" {

typedef struct oracle_sim oracle_sim;

const char* oracle_last_error(void) { return g_err.c_str(); }

oracle_sim* oracle_create(int precision, const asph_parameters* p, const asph_vec4* positions, uint32_t count,
                          const asph_boundary_particle* boundary, uint32_t nb, const asph_refinement_parameters* r,
                          uint32_t maxFluid, uint32_t maxBoundary)
{
    try
    {
        auto* h = new OracleHandle();
        h->precision = precision;
        if (precision == 64)
        {
            h->d = makeSim<double>(*p, positions, count, boundary, nb, r, maxFluid, maxBoundary);
        }
        else
        {
            h->f = makeSim<float>(*p, positions, count, boundary, nb, r, maxFluid, maxBoundary);
        }
        return reinterpret_cast<oracle_sim*>(h);
    }
    catch (const std::exception& e)
    {
        g_err = e.what();
        return nullptr;
    }
}

void oracle_destroy(oracle_sim* s) { delete reinterpret_cast<OracleHandle*>(s); }

#define ORACLE_DISPATCH(handle, ...)                         \
    do                                                       \
    {                                                        \
        auto* H_ = reinterpret_cast<OracleHandle*>(handle);  \
        if (H_->precision == 64)                             \
        {                                                    \
            auto& S = *H_->d;                                \
            __VA_ARGS__;                                     \
        }                                                    \
        else                                                 \
        {                                                    \
            auto& S = *H_->f;                                \
            __VA_ARGS__;                                     \
        }                                                    \
    } while (0)

void oracle_update(oracle_sim* s, float dt) { ORACLE_DISPATCH(s, S.update(dt)); }

uint32_t oracle_count(oracle_sim* s)
{
    uint32_t c = 0;
    ORACLE_DISPATCH(s, c = S.n);
    return c;
}

void oracle_enable_refinement(oracle_sim* s) { ORACLE_DISPATCH(s, S.frame = S.ref.initialCooldown); }
void oracle_set_frame(oracle_sim* s, uint32_t f) { ORACLE_DISPATCH(s, S.frame = f); }

// Fluid state as float (rounded when the oracle runs in FP64).
void oracle_copy_fluid(oracle_sim* s, float* pos4, float* vel4, float* acc4, float* rho, float* rad, float* h, float* m)
{
    ORACLE_DISPATCH(s, {
        for (uint32_t i = 0; i < S.n; ++i)
        {
            if (pos4 != nullptr)
            {
                pos4[4 * i] = float(S.pos[i].x);
                pos4[4 * i + 1] = float(S.pos[i].y);
                pos4[4 * i + 2] = float(S.pos[i].z);
                pos4[4 * i + 3] = float(S.pos[i].w);
            }
            if (vel4 != nullptr)
            {
                vel4[4 * i] = float(S.vel[i].x);
                vel4[4 * i + 1] = float(S.vel[i].y);
                vel4[4 * i + 2] = float(S.vel[i].z);
                vel4[4 * i + 3] = float(S.vel[i].w);
            }
            if (acc4 != nullptr)
            {
                acc4[4 * i] = float(S.acc[i].x);
                acc4[4 * i + 1] = float(S.acc[i].y);
                acc4[4 * i + 2] = float(S.acc[i].z);
                acc4[4 * i + 3] = float(S.acc[i].w);
            }
            if (rho != nullptr) rho[i] = float(S.rho[i]);
            if (rad != nullptr) rad[i] = float(S.rad[i]);
            if (h != nullptr) h[i] = float(S.h[i]);
            if (m != nullptr) m[i] = float(S.m[i]);
        }
    });
}

void oracle_copy_fluid64(oracle_sim* s, double* pos4, double* vel4, double* acc4, double* rho)
{
    ORACLE_DISPATCH(s, {
        for (uint32_t i = 0; i < S.n; ++i)
        {
            if (pos4 != nullptr)
            {
                pos4[4 * i] = S.pos[i].x;
                pos4[4 * i + 1] = S.pos[i].y;
                pos4[4 * i + 2] = S.pos[i].z;
                pos4[4 * i + 3] = S.pos[i].w;
            }
            if (vel4 != nullptr)
            {
                vel4[4 * i] = S.vel[i].x;
                vel4[4 * i + 1] = S.vel[i].y;
                vel4[4 * i + 2] = S.vel[i].z;
                vel4[4 * i + 3] = S.vel[i].w;
            }
            if (acc4 != nullptr)
            {
                acc4[4 * i] = S.acc[i].x;
                acc4[4 * i + 1] = S.acc[i].y;
                acc4[4 * i + 2] = S.acc[i].z;
                acc4[4 * i + 3] = S.acc[i].w;
            }
            if (rho != nullptr) rho[i] = S.rho[i];
        }
    });
}

// Replace the fluid state (float inputs).
int oracle_set_fluid(oracle_sim* s, uint32_t count, const float* pos4, const float* vel4, const float* acc4,
                     const float* rho, const float* rad, const float* h, const float* m)
{
    int rc = 0;
    ORACLE_DISPATCH(s, {
        if (count > S.cap)
        {
            rc = 1;
        }
        else
        {
            S.n = count;
            for (uint32_t i = 0; i < count; ++i)
            {
                if (pos4 != nullptr) S.pos[i] = {pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], pos4[4 * i + 3]};
                if (vel4 != nullptr) S.vel[i] = {vel4[4 * i], vel4[4 * i + 1], vel4[4 * i + 2], vel4[4 * i + 3]};
                if (acc4 != nullptr) S.acc[i] = {acc4[4 * i], acc4[4 * i + 1], acc4[4 * i + 2], acc4[4 * i + 3]};
                if (rho != nullptr) S.rho[i] = rho[i];
                if (rad != nullptr) S.rad[i] = rad[i];
                if (h != nullptr) S.h[i] = h[i];
                if (m != nullptr) S.m[i] = m[i];
            }
        }
    });
    return rc;
}

void oracle_grid_dims(oracle_sim* s, int32_t* dims)
{
    ORACLE_DISPATCH(s, {
        dims[0] = S.gx;
        dims[1] = S.gy;
        dims[2] = S.gz;
    });
}

// Rebuild the fluid grid from the current positions (the grid of the last
// update is what oracle_grid returns otherwise).
void oracle_rebuild_grid(oracle_sim* s) { ORACLE_DISPATCH(s, S.buildGrid(S.fluid, S.pos, S.n)); }

// Number of particles in the fluid grid (the count when it was built; a
// refinement step changes the count afterwards).
uint32_t oracle_grid_count(oracle_sim* s)
{
    uint32_t n = 0;
    ORACLE_DISPATCH(s, n = static_cast<uint32_t>(S.fluid.keys.size()));
    return n;
}

void oracle_grid(oracle_sim* s, int32_t* keys, int32_t* idx, int32_t* cellStart, int32_t* cellEnd)
{
    ORACLE_DISPATCH(s, {
        const auto& g = S.fluid;
        if (keys != nullptr) std::copy(g.keys.begin(), g.keys.end(), keys);
        if (idx != nullptr) std::copy(g.idx.begin(), g.idx.end(), idx);
        if (cellStart != nullptr) std::copy(g.cellStart.begin(), g.cellStart.end(), cellStart);
        if (cellEnd != nullptr) std::copy(g.cellEnd.begin(), g.cellEnd.end(), cellEnd);
    });
}

void oracle_boundary_grid(oracle_sim* s, int32_t* keys, int32_t* idx, int32_t* cellStart, int32_t* cellEnd)
{
    ORACLE_DISPATCH(s, {
        const auto& g = S.boundary;
        if (keys != nullptr) std::copy(g.keys.begin(), g.keys.end(), keys);
        if (idx != nullptr) std::copy(g.idx.begin(), g.idx.end(), idx);
        if (cellStart != nullptr) std::copy(g.cellStart.begin(), g.cellStart.end(), cellStart);
        if (cellEnd != nullptr) std::copy(g.cellEnd.begin(), g.cellEnd.end(), cellEnd);
    });
}

// Neighbour sets over the current grid and positions.
void oracle_neighbor_sets(oracle_sim* s, uint32_t* counts, uint64_t* prints)
{
    ORACLE_DISPATCH(s, S.neighborSets(counts, prints));
}

// calculateAverageNeighborCount (SphSimulation.cu:237-247)
float oracle_average_neighbor_count(oracle_sim* s)
{
    float r = 0.F;
    ORACLE_DISPATCH(s, {
        std::vector<uint32_t> c(S.n);
        S.neighborSets(c.data(), nullptr);
        uint32_t total = 0;
        for (const uint32_t v : c) total += v;
        r = S.n > 0 ? static_cast<float>(total) / static_cast<float>(S.n) : 0.F;
    });
    return r;
}

// getDensityInfo (SphSimulation.cu:249-285); the sum is accumulated in FP64
// (thrust's FP32 reduction order is unpinned).
void oracle_density_info(oracle_sim* s, float threshold, asph_density_info* out)
{
    ORACLE_DISPATCH(s, {
        const float rho0 = float(S.p.rho0);
        asph_density_info info {rho0, FLT_MAX, -FLT_MAX, 0.F, 0U, 0U, 0U};
        double sum = 0.0;
        for (uint32_t i = 0; i < S.n; ++i)
        {
            const float d = float(S.rho[i]);
            const float ratio = d / rho0 - 1.0F;
            info.min_density = std::min(info.min_density, d);
            info.max_density = std::max(info.max_density, d);
            sum += d;
            info.under_density_count += (ratio < -threshold) ? 1U : 0U;
            info.normal_density_count += (ratio >= -threshold && ratio <= threshold) ? 1U : 0U;
            info.over_density_count += (ratio > threshold) ? 1U : 0U;
        }
        info.average_density = static_cast<float>(sum / static_cast<double>(S.n));
        *out = info;
    });
}

void oracle_refinement_trace(oracle_sim* s, asph_refinement_trace* t, uint32_t* mergeList, uint32_t* cand,
                             uint32_t* pairKeep, uint32_t* splitParents)
{
    ORACLE_DISPATCH(s, {
        const auto& tr = S.trace;
        *t = tr.t;
        if (mergeList != nullptr) std::copy(tr.mergeList.begin(), tr.mergeList.end(), mergeList);
        if (cand != nullptr) std::copy(tr.candidates.begin(), tr.candidates.end(), cand);
        if (pairKeep != nullptr) std::copy(tr.pairKeep.begin(), tr.pairKeep.end(), pairKeep);
        if (splitParents != nullptr) std::copy(tr.splitParents.begin(), tr.splitParents.end(), splitParents);
    });
}

// Split particle `parent` into slots [base, base + 12) of the current state
// (splitParticles body, AdaptiveAlgorithm.cu:88-116); count grows to base + 12.
int oracle_split_one(oracle_sim* s, uint32_t parent, uint32_t base)
{
    int rc = 0;
    ORACLE_DISPATCH(s, {
        if (parent >= S.n || base + 12U > S.cap)
        {
            rc = 1;
        }
        else
        {
            S.splitOne(parent, base);
            S.n = std::max(S.n, base + 12U);
        }
    });
    return rc;
}

// Merged smoothing length of two particles (calculateMergedSmoothingLength).
float oracle_merged_h(const float* pa, const float* pb, float ma, float mb, float ha, float hb, const float* pm,
                      float M, uint32_t flags)
{
    oracle::Sim<float> tmp;
    tmp.p.flags = flags;
    const oracle::V4<float> a {pa[0], pa[1], pa[2], pa[3]};
    const oracle::V4<float> b {pb[0], pb[1], pb[2], pb[3]};
    const oracle::V4<float> m {pm[0], pm[1], pm[2], pm[3]};
    return tmp.mergedSmoothingLength(a, b, ma, mb, ha, hb, m, M);
}

// findTopParticlesToSplit / ToMerge selection on a value array.
uint32_t oracle_top_list(const float* values, uint32_t n, int descending, float ratio, uint32_t* out)
{
    oracle::Sim<float> tmp;
    tmp.ref.batchRatio = ratio;
    const std::vector<float> v(values, values + n);
    uint32_t k = 0;
    const auto ids = tmp.topList(v, descending != 0, &k);
    std::copy(ids.begin(), ids.end(), out);
    return k;
}

// calculateCellStartAndEndIndices on given sorted keys (cells entries each).
void oracle_cell_ranges(const int32_t* sortedKeys, uint32_t n, uint32_t cells, uint32_t fixQuirk, int32_t* start,
                        int32_t* end)
{
    oracle::Sim<float> tmp;
    tmp.p.flags = fixQuirk != 0U ? ASPH_FLAG_FIX_CELL_START : 0U;
    tmp.gx = static_cast<int>(cells);
    tmp.gy = 1;
    tmp.gz = 1;
    // reuse buildGrid's range rule on keys already sorted
    oracle::Grid g;
    g.count = n;
    g.cellStart.assign(cells, -1);
    g.cellEnd.assign(cells, -1);
    g.keys.assign(sortedKeys, sortedKeys + n);
    const bool fix = fixQuirk != 0U;
    for (uint32_t sl = 0; sl < n; ++sl)
    {
        const int32_t c = g.keys[sl];
        if (sl == n - 1)
        {
            g.cellEnd[c] = static_cast<int32_t>(sl);
            if (fix && sl == 0)
            {
                g.cellStart[c] = 0;
            }
            continue;
        }
        if (sl == 0)
        {
            g.cellStart[c] = 0;
            if (!fix)
            {
                continue;
            }
        }
        const int32_t cn = g.keys[sl + 1];
        if (c != cn)
        {
            g.cellStart[cn] = static_cast<int32_t>(sl + 1);
            g.cellEnd[c] = static_cast<int32_t>(sl);
        }
    }
    std::copy(g.cellStart.begin(), g.cellStart.end(), start);
    std::copy(g.cellEnd.begin(), g.cellEnd.end(), end);
}

// Criterion values of every particle of the current state (phase 0: merge
// generators, 1: split generators; getCriterionValues, AdaptiveAlgorithm.cuh:
// 40-54) over a grid rebuilt from the current positions, with the zone
// centre at time t.
void oracle_criteria(oracle_sim* s, int phase, double t, float* out)
{
    ORACLE_DISPATCH(s, {
        S.buildGrid(S.fluid, S.pos, S.n);
        const auto zc = S.zoneCenterAt(t);
        for (uint32_t i = 0; i < S.n; ++i)
        {
            out[i] = static_cast<float>(phase == 0 ? S.mergeCriterion(i, S.fluid, zc) : S.splitCriterion(i, S.fluid, zc));
        }
    });
}

// S_i of every particle (Sim::forceScale) over the grid of the last update.
void oracle_force_scale(oracle_sim* s, double* out) { ORACLE_DISPATCH(s, S.forceScale(out)); }

// OpenMP threads the oracle's parallel loops use.
int oracle_threads(void) { return omp_get_max_threads(); }

// Scalar closed forms for the math pins.
float oracle_wendland_f32(float d, float h) { return oracle::wendland<float>(d, h); }
double oracle_wendland_f64(double d, double h) { return oracle::wendland<double>(d, h); }
float oracle_wendland_grad_f32(float d, float h) { return oracle::wendlandDerivative<float>(d, h); }
double oracle_wendland_grad_f64(double d, double h) { return oracle::wendlandDerivative<double>(d, h); }
float oracle_tait_f32(float rho, float rho0, float c, float g) { return oracle::taitPressure<float>(rho, rho0, c, g); }
float oracle_powf(float x, float g) { return oracle::det_powf(x, g); }
float oracle_cbrtf(float x) { return oracle::det_cbrtf(x); }
void oracle_sincos(double a, double* s, double* c) { oracle::det_sincos(a, s, c); }
uint64_t oracle_mix64(uint64_t z) { return oracle::mix64(z); }

// StaticBoundaryDomain::generate restated (wall lattice + psi).
uint32_t oracle_boundary_generate(const asph_domain* d, float spacing, float rho0, float h, asph_boundary_particle* out)
{
    const auto pts = oracle::wallParticles(*d, spacing);
    if (out == nullptr)
    {
        return static_cast<uint32_t>(pts.size());
    }
    const auto psi = oracle::wallPsi(pts, h, rho0);
    for (size_t i = 0; i < pts.size(); ++i)
    {
        out[i].position = {pts[i].x, pts[i].y, pts[i].z, pts[i].w};
        out[i].psi = psi[i];
    }
    return static_cast<uint32_t>(pts.size());
}

}  // extern "C"
