// Host side of the HIP engine (see engine.hpp).
#include "engine.hpp"

#include "walls.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

namespace asph
{

void hipCheck(hipError_t e, const char* what)
{
    if (e != hipSuccess)
    {
        throw Error(ASPH_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// ------------------------------------------------------------ allocator --

RangeAllocator::RangeAllocator(uint64_t capacity) : cap_(capacity)
{
    if (capacity > 0)
    {
        free_.push_back({0, capacity});
    }
}

uint64_t RangeAllocator::alloc(uint64_t n)
{
    if (n == 0)
    {
        return 0;
    }
    for (size_t k = 0; k < free_.size(); ++k)
    {
        if (free_[k].n >= n)
        {
            const uint64_t off = free_[k].off;
            free_[k].off += n;
            free_[k].n -= n;
            if (free_[k].n == 0)
            {
                free_.erase(free_.begin() + static_cast<long>(k));
            }
            used_ += n;
            return off;
        }
    }
    return UINT64_MAX;
}

bool RangeAllocator::claim(uint64_t off, uint64_t n)
{
    if (n == 0)
    {
        return true;
    }
    for (size_t k = 0; k < free_.size(); ++k)
    {
        const Range r = free_[k];
        if (r.off <= off && off + n <= r.off + r.n)
        {
            free_.erase(free_.begin() + static_cast<long>(k));
            if (off + n < r.off + r.n)
            {
                free_.insert(free_.begin() + static_cast<long>(k), Range {off + n, r.off + r.n - (off + n)});
            }
            if (r.off < off)
            {
                free_.insert(free_.begin() + static_cast<long>(k), Range {r.off, off - r.off});
            }
            used_ += n;
            return true;
        }
    }
    return false;
}

void RangeAllocator::grow(uint64_t capacity)
{
    if (capacity <= cap_)
    {
        return;
    }
    const uint64_t extra = capacity - cap_;
    const uint64_t off = cap_;
    cap_ = capacity;
    used_ += extra;  // release() subtracts it again
    release(off, extra);
}

void RangeAllocator::release(uint64_t off, uint64_t n)
{
    if (n == 0)
    {
        return;
    }
    used_ -= n;
    free_.push_back({off, n});
    std::sort(free_.begin(), free_.end(), [](const Range& a, const Range& b) { return a.off < b.off; });
    std::vector<Range> merged;
    for (const auto& r : free_)
    {
        if (!merged.empty() && merged.back().off + merged.back().n == r.off)
        {
            merged.back().n += r.n;
        }
        else
        {
            merged.push_back(r);
        }
    }
    free_ = merged;
}

// --------------------------------------------------------------- grid --

// NeighborGrid::createGrid (NeighborGrid.cu:108-121) with the cell width of
// SphSimulation.cu:46-49 (2 * h0 * wendlandRangeRatio), in FP32 on the host.
GridGeom makeGrid(const asph_domain& d, float h0)
{
    const float cellWidth = (2.0F * h0) * 2.0F;
    const float sx = d.max.x - d.min.x;
    const float sy = d.max.y - d.min.y;
    const float sz = d.max.z - d.min.z;
    const float n[3] = {std::ceil(sx / cellWidth), std::ceil(sy / cellWidth), std::ceil(sz / cellWidth)};
    double cells = 1.0;
    for (const float v : n)
    {
        // also catches NaN (an empty axis, a non-positive or non-finite h0)
        if (!(v >= 1.0F) || !(v <= static_cast<float>(kMaxGridCells)))
        {
            throw Error(ASPH_ERR_INVALID, "neighbour grid would be empty or need more than 1e9 cells (domain "
                                          "empty on an axis, or smoothing radius too small for the domain)");
        }
        cells *= static_cast<double>(v);
    }
    if (cells > static_cast<double>(kMaxGridCells))
    {
        throw Error(ASPH_ERR_INVALID, "neighbour grid would need more than 1e9 cells (smoothing radius too small "
                                      "for the domain)");
    }
    GridGeom g;
    g.gx = static_cast<int>(n[0]);
    g.gy = static_cast<int>(n[1]);
    g.gz = static_cast<int>(n[2]);
    g.cs[0] = sx / static_cast<float>(g.gx);
    g.cs[1] = sy / static_cast<float>(g.gy);
    g.cs[2] = sz / static_cast<float>(g.gz);
    return g;
}

namespace
{

inline float hmin(float a, float b) { return (b < a) ? b : a; }
inline float hmax(float a, float b) { return (a < b) ? b : a; }

// |v| in [2^-40, 2^39]: a divisor the device may divide by through its
// reciprocal with the quotient's exact bits (kDivLo, kDivHi of
// device_common.hpp bound that at [2^-40, 2^40]), and a smoothing length whose
// support (2 h)^2 stays within 2^80, so that the uniform density loop's
// distances lie in the domain of its square root and quotient
inline bool sharedDivisor(float v)
{
    uint32_t u;
    std::memcpy(&u, &v, sizeof(u));
    return ((u & 0x7FFFFFFFU) - 0x2B800000U) <= 0x27800000U;
}

// calculateCellIndex + flattenCellIndex (NeighborGrid.cuh:166-177) on the host.
int32_t hostCellKey(const SceneDev& sc, float px, float py, float pz)
{
    const float scx = sc.dmax[0] - sc.dmin[0];
    const float scy = sc.dmax[1] - sc.dmin[1];
    const float scz = sc.dmax[2] - sc.dmin[2];
    const float rx = hmin(hmax(px - sc.dmin[0], 0.0F), scx);
    const float ry = hmin(hmax(py - sc.dmin[1], 0.0F), scy);
    const float rz = hmin(hmax(pz - sc.dmin[2], 0.0F), scz);
    const int cx = std::min(static_cast<int>(rx / sc.cs[0]), sc.gx - 1);
    const int cy = std::min(static_cast<int>(ry / sc.cs[1]), sc.gy - 1);
    const int cz = std::min(static_cast<int>(rz / sc.cs[2]), sc.gz - 1);
    return cx + (cy * sc.gx) + (cz * sc.gx * sc.gy);
}

int bitsFor(uint64_t maxValue)
{
    int b = 1;
    while (b < 63 && (1ULL << b) <= maxValue)
    {
        ++b;
    }
    return b;
}

}  // namespace

void validateParams(const asph_parameters& p)
{
    auto bad = [](const char* m) { throw Error(ASPH_ERR_INVALID, m); };
    if (!(p.base_smoothing_radius > 0.0F) || !std::isfinite(p.base_smoothing_radius))
    {
        bad("base_smoothing_radius must be positive and finite");
    }
    if (!(p.domain.max.x > p.domain.min.x) || !(p.domain.max.y > p.domain.min.y) ||
        !(p.domain.max.z > p.domain.min.z))
    {
        bad("domain.max must exceed domain.min on every axis");
    }
    if (!(p.rest_density > 0.0F))
    {
        bad("rest_density must be positive");
    }
    if (!(p.laminar_viscosity >= 0.0F) || !std::isfinite(p.laminar_viscosity))
    {
        bad("laminar_viscosity must be finite and non-negative");
    }
    if (!(p.tait_gamma > 0.0F))
    {
        bad("tait_gamma must be positive");
    }
    if (p.threads_per_block == 0)
    {
        bad("threads_per_block must be positive");
    }
    (void)makeGrid(p.domain, p.base_smoothing_radius);  // throws for grids beyond kMaxGridCells
}

uint32_t listCapacityRounded(uint32_t listCap)
{
    const uint32_t k = listCap == 0U ? kListRowEntries : listCap;
    return ((k + kListRowEntries - 1U) / kListRowEntries) * kListRowEntries;
}

uint64_t sceneBytes(uint32_t capacity, uint32_t boundaryCount, uint32_t cells, bool adaptive, uint32_t listCap)
{
    // Per fluid slot: canonical arrays 64, sort keys/values 16, sorted shadows
    // s0/s1/fa/fb 64 + shm 16 + scene id 4 + neighbour count 4, group boxes 4,
    // neighbour list 4 per entry (capacity rounded up to whole 4-entry
    // chunks), overflow and wall queues 8, query order 8, radix histogram and
    // scan tiles < 1.
    uint64_t perParticle = 64ULL + 16ULL + 88ULL + 4ULL + 4ULL * listCapacityRounded(listCap) + 8ULL + 8ULL + 1ULL;
    if (adaptive)
    {
        // criterion keys/values (two buffers) 24, candidate/removal/keep/prefix
        // 16, compaction copies 64, merge list capture 4, sorted slot 4
        perParticle += 24ULL + 16ULL + 64ULL + 4ULL + 4ULL;
    }
    // boundary slot: position + psi 16, viscosity 4, chunk box 32; cell: count,
    // offset and boundary offset 12, boundary box 32, wall distance bound 4.
    return perParticle * capacity + 52ULL * boundaryCount + 48ULL * (cells + 1ULL);
}

// --------------------------------------------------------------- engine --

template <class T>
T* Engine::alloc(size_t n)
{
    n = std::max<size_t>(n, 1);
    void* p = nullptr;
    hipCheck(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
    allocations_.push_back(p);
    bytes_ += n * sizeof(T);
    return static_cast<T*>(p);
}

Engine::Engine(const EngineConfig& cfg)
    : cfg_(cfg), pAlloc_(cfg.particleCapacity), bAlloc_(cfg.boundaryCapacity), cAlloc_(cfg.cellCapacity)
{
    if (cfg_.sceneSlots == 0)
    {
        throw Error(ASPH_ERR_INVALID, "sceneSlots must be positive");
    }
    if (cfg_.particleCapacity > kMaxParticleCapacity)
    {
        // the neighbour loops address 16-byte particle records by 32-bit byte offsets
        throw Error(ASPH_ERR_CAPACITY, "particle capacity above 2^28 - 64");
    }
    hipCheck(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    try
    {
        const uint32_t P = std::max<uint32_t>(cfg_.particleCapacity, 1);
        const uint32_t B = std::max<uint32_t>(cfg_.boundaryCapacity, 1);
        const uint32_t C = std::max<uint32_t>(cfg_.cellCapacity, 2);
        const uint32_t S = cfg_.sceneSlots;
        Ctx& c = ctx_;
        c.scenes = alloc<SceneDev>(S);
        c.states = alloc<SceneState>(S);
        c.sceneSlots = S;
        c.denseStarts = alloc<uint32_t>(S);
        c.sortedStarts = alloc<uint32_t>(S);
        cellOrderDev_ = alloc<uint32_t>(S);
        c.cellOrder = cellOrderDev_;
        cellBaseOrderDev_ = alloc<uint32_t>(S);
        c.cellBaseOrder = cellBaseOrderDev_;
        nOrderDev_ = alloc<uint32_t>(1);
        c.nOrderPtr = nOrderDev_;
        c.nTotal = alloc<uint32_t>(1);
        c.pos = cfg_.importedPos != nullptr ? cfg_.importedPos : alloc<F4>(P);
        c.vel = cfg_.importedVel != nullptr ? cfg_.importedVel : alloc<F4>(P);
        c.acc = alloc<F4>(P);
        c.rho = cfg_.importedRho != nullptr ? cfg_.importedRho : alloc<float>(P);
        c.rad = cfg_.importedRad != nullptr ? cfg_.importedRad : alloc<float>(P);
        c.h = alloc<float>(P);
        c.m = alloc<float>(P);
        c.pCap = P;
        c.keysA = alloc<uint32_t>(P);
        c.valsA = alloc<uint32_t>(P);
        c.keysB = alloc<uint32_t>(P);
        c.valsB = alloc<uint32_t>(P);
        c.cellOff = alloc<uint32_t>(static_cast<size_t>(C) + 1);
        c.cCap = C;
        c.s0 = alloc<F4>(P);
        c.s1 = alloc<F4>(P);
        c.shm = alloc<F4>(P);
        c.fa = alloc<F4>(P);
        c.fb = alloc<F4>(P);
        c.gbox0 = alloc<F4>(P / kGroup + 1);
        c.gbox1 = alloc<F4>(P / kGroup + 1);
        c.sscene = alloc<uint32_t>(P);
        c.ncount = alloc<uint32_t>(P);
        c.overflowList = alloc<uint32_t>(P);
        c.overflowCount = alloc<uint32_t>(1);
        c.wallQueue = alloc<uint32_t>(P);
        c.wallQueueCount = alloc<uint32_t>(1);
        c.queryAny = alloc<uint32_t>(1);
        c.wcost = alloc<uint32_t>((static_cast<size_t>(P) + 63U) / 64U + 1U);
        c.wheavy = alloc<uint32_t>((static_cast<size_t>(P) + 63U) / 64U + 1U);
        c.heavyList = alloc<uint32_t>(kHeavyCap);
        c.heavyCount = alloc<uint32_t>(1);
        c.wstat = alloc<uint32_t>(4);
        cfg_.listCap = listCapacityRounded(cfg_.listCap);
        c.listCap = cfg_.listCap;
        // wave-blocked rows: whole waves of 64 slots
        c.list = alloc<uint32_t>(static_cast<size_t>(cfg_.listCap) * ((static_cast<size_t>(P) + 63U) / 64U * 64U));
        c.bs0 = alloc<F4>(B);
        c.bvisc = alloc<float>(B);
        c.bcellOff = alloc<uint32_t>(static_cast<size_t>(C) + 1);
        c.bbox0 = alloc<F4>(static_cast<size_t>(C) + 1);
        c.bbox1 = alloc<F4>(static_cast<size_t>(C) + 1);
        c.bnear2 = alloc<float>(static_cast<size_t>(C) + 1);
        c.bgbox0 = alloc<F4>(B);
        c.bgbox1 = alloc<F4>(B);
        qpermBuf_ = alloc<uint32_t>(P);
        qposBuf_ = alloc<uint32_t>(P);
        int cus = 0;
        int dev = 0;
        hipCheck(hipGetDevice(&dev), "hipGetDevice");
        hipCheck(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        computeUnits_ = static_cast<uint32_t>(std::max(cus, 1));
        if (cfg_.adaptive)
        {
            c.ckeysA = alloc<uint64_t>(P);
            c.cvalsA = alloc<uint32_t>(P);
            c.ckeysB = alloc<uint64_t>(P);
            c.cvalsB = alloc<uint32_t>(P);
            c.cand = alloc<uint32_t>(P);
            c.slotOf = alloc<uint32_t>(P);
            c.critCount = alloc<uint32_t>(1);
            c.removal = alloc<uint32_t>(P);
            c.keep = alloc<uint32_t>(P);
            c.rprefix = alloc<uint32_t>(P);
            c.xpos = alloc<F4>(P);
            c.xvel = alloc<F4>(P);
            c.xacc = alloc<F4>(P);
            c.xrho = alloc<float>(P);
            c.xrad = alloc<float>(P);
            c.xh = alloc<float>(P);
            c.xm = alloc<float>(P);
            c.mergeListDbg = alloc<uint32_t>(P);
        }
        hist_ = alloc<uint32_t>(radixHistWords(P, kMaxRadixBits));
        digitTotals_ = alloc<uint32_t>(1U << kMaxRadixBits);
        tileSums_ = alloc<uint32_t>(scanTileWords(std::max<uint32_t>(P, C + 1)));
        scanTotal_ = alloc<uint32_t>(1);
        infoD_ = alloc<double>(kInfoBlocks);
        infoU_ = alloc<uint32_t>(5 * kInfoBlocks);
        diagTotal_ = alloc<uint32_t>(1);

        hipCheck(hipMemset(c.states, 0, sizeof(SceneState) * S), "hipMemset");
        hipCheck(hipMemset(const_cast<SceneDev*>(c.scenes), 0, sizeof(SceneDev) * S), "hipMemset");
        hipCheck(hipMemset(c.cellOff, 0, sizeof(uint32_t) * (C + 1)), "hipMemset");
        hipCheck(hipMemset(c.bcellOff, 0, sizeof(uint32_t) * (C + 1)), "hipMemset");
        hipCheck(hipMemset(nOrderDev_, 0, sizeof(uint32_t)), "hipMemset");
        hipCheck(hipMemset(c.nTotal, 0, sizeof(uint32_t)), "hipMemset");
        hipCheck(hipMemset(c.queryAny, 0, sizeof(uint32_t)), "hipMemset");
        hipCheck(hipMemset(c.wcost, 0, sizeof(uint32_t) * ((static_cast<size_t>(P) + 63U) / 64U + 1U)), "hipMemset");
        hipCheck(hipMemset(c.wheavy, 0, sizeof(uint32_t) * ((static_cast<size_t>(P) + 63U) / 64U + 1U)), "hipMemset");
        hipCheck(hipMemset(c.heavyCount, 0, sizeof(uint32_t)), "hipMemset");
        hipCheck(hipMemset(c.wstat, 0, sizeof(uint32_t) * 4), "hipMemset");
        if (cfg_.adaptive)
        {
            hipCheck(hipMemset(c.removal, 0, sizeof(uint32_t) * P), "hipMemset");
            hipCheck(hipMemset(c.keep, 0, sizeof(uint32_t) * P), "hipMemset");
            hipCheck(hipMemset(c.cand, 0, sizeof(uint32_t) * P), "hipMemset");
        }
        hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
        // criterion keys: scene << 32 | 32 ordered value bits of the
        // participating scenes only, sorted on 32 + sceneBits_ bits
        sceneBits_ = S <= 1U ? 0 : bitsFor(static_cast<uint64_t>(S) - 1U);
        scenes_.resize(S);
        c.nMax = 1;
        c.cExtent = 2;
        c.forceKinds = kForceExact;
        keyBits_ = 1;
        buildCtx();
    }
    catch (...)
    {
        for (void* p : allocations_)
        {
            (void)hipFree(p);
        }
        allocations_.clear();
        (void)hipStreamDestroy(stream_);
        throw;
    }
}

Engine::~Engine()
{
    (void)hipStreamSynchronize(stream_);
    for (auto& marks : pendingMarks_)
    {
        for (hipEvent_t e : marks)
        {
            (void)hipEventDestroy(e);
        }
    }
    invalidateGraphs();
    for (void* p : allocations_)
    {
        (void)hipFree(p);
    }
    (void)hipStreamDestroy(stream_);
}

void Engine::invalidateGraphs()
{
    if (graphPlain_ != nullptr)
    {
        (void)hipGraphExecDestroy(graphPlain_);
        graphPlain_ = nullptr;
    }
    if (graphRefine_ != nullptr)
    {
        (void)hipGraphExecDestroy(graphRefine_);
        graphRefine_ = nullptr;
    }
    if (graphPlainG_ != nullptr)
    {
        (void)hipGraphDestroy(graphPlainG_);
        graphPlainG_ = nullptr;
    }
    if (graphRefineG_ != nullptr)
    {
        (void)hipGraphDestroy(graphRefineG_);
        graphRefineG_ = nullptr;
    }
}

void Engine::buildCtx()
{
    Ctx& c = ctx_;
    const int tile = cfg_.launch.queryTile < 0 ? static_cast<int>(kQueryAuto) : cfg_.launch.queryTile;
    c.qperm = tile > 0 ? qpermBuf_ : nullptr;
    c.qpos = tile > 0 ? qposBuf_ : nullptr;
    c.queryTileBits = static_cast<uint32_t>(tile);
    cfg_.launch.forceGrid = cfg_.launch.forceBlocksPerCU * computeUnits_;
    const int rb = radixDigitBits(cfg_.launch.radixBits, keyBits_);
    const int passes = (keyBits_ + rb - 1) / rb;
    sortParity_ = passes & 1;
    c.sortedKeys = sortParity_ != 0 ? c.keysB : c.keysA;
    c.sortedVals = sortParity_ != 0 ? c.valsB : c.valsA;
    const int passes64 = (32 + sceneBits_ + 7) / 8;
    csortParity_ = passes64 & 1;
    if (cfg_.adaptive)
    {
        c.csortedKeys = csortParity_ != 0 ? c.ckeysB : c.ckeysA;
        c.csortedVals = csortParity_ != 0 ? c.cvalsB : c.cvalsA;
    }
}

void Engine::setLaunch(const LaunchCfg& lc)
{
    synchronize();
    cfg_.launch = lc;
    buildCtx();
    invalidateGraphs();
}

void Engine::setUseGraph(bool on)
{
    synchronize();
    cfg_.useGraph = on;
    invalidateGraphs();
}

uint32_t Engine::addScene(const asph_parameters& p, const asph_vec4* positions, uint32_t count,
                          const asph_boundary_particle* boundary, uint32_t nb, const asph_refinement_parameters* r,
                          uint32_t capacity, float dt)
{
    validateParams(p);
    const bool adaptive = r != nullptr && r->enabled != 0U;
    if (adaptive)
    {
        if (!cfg_.adaptive)
        {
            throw Error(ASPH_ERR_STATE, "engine was created without refinement buffers");
        }
        if (r->cooldown == 0U)
        {
            throw Error(ASPH_ERR_INVALID, "refinement cooldown must be positive");
        }
        if (r->criterion_type > ASPH_CRITERION_ZONE)
        {
            throw Error(ASPH_ERR_INVALID, "unknown refinement criterion");
        }
        capacity = r->max_particle_count;  // AdaptiveSphSimulation.cu:37
    }
    else if (capacity == 0)
    {
        capacity = count;
    }
    if (count > capacity)
    {
        throw Error(ASPH_ERR_CAPACITY, "particle count exceeds the fluid capacity");
    }
    if (count > 0 && positions == nullptr)
    {
        throw Error(ASPH_ERR_INVALID, "positions is NULL");
    }
    if (nb > 0 && boundary == nullptr)
    {
        throw Error(ASPH_ERR_INVALID, "boundary is NULL");
    }
    for (uint32_t i = 0; i < count; ++i)
    {
        const auto& q = positions[i];
        if (!std::isfinite(q.x) || !std::isfinite(q.y) || !std::isfinite(q.z) || q.w != 0.0F)
        {
            throw Error(ASPH_ERR_INVALID, "positions must be finite with w == 0");
        }
    }
    uint32_t id = UINT32_MAX;
    for (uint32_t s = 0; s < scenes_.size(); ++s)
    {
        if (!scenes_[s].active)
        {
            id = s;
            break;
        }
    }
    if (id == UINT32_MAX)
    {
        throw Error(ASPH_ERR_CAPACITY, "no free scene slot");
    }
    const GridGeom g = makeGrid(p.domain, p.base_smoothing_radius);
    const uint64_t cells = g.cells();
    if (g.gx <= 0 || g.gy <= 0 || g.gz <= 0 || cells > 0x7FFFFFFFULL)
    {
        throw Error(ASPH_ERR_INVALID, "grid is empty or too large");
    }
    const uint64_t pOff = pAlloc_.alloc(capacity);
    if (pOff == UINT64_MAX)
    {
        throw Error(ASPH_ERR_CAPACITY, "particle pool exhausted");
    }
    const uint64_t bOff = bAlloc_.alloc(nb);
    if (bOff == UINT64_MAX)
    {
        pAlloc_.release(pOff, capacity);
        throw Error(ASPH_ERR_CAPACITY, "boundary pool exhausted");
    }
    const uint64_t cBase = cAlloc_.alloc(cells + 1);
    if (cBase == UINT64_MAX)
    {
        pAlloc_.release(pOff, capacity);
        bAlloc_.release(bOff, nb);
        throw Error(ASPH_ERR_CAPACITY, "cell pool exhausted");
    }
    synchronize();
    HostScene& hs = scenes_[id];
    hs = HostScene {};
    hs.active = true;
    hs.adaptive = adaptive;
    hs.cellRegion = static_cast<uint32_t>(cells + 1);
    hs.count0 = count;
    hs.knownH = count;
    hs.knownM = count;
    SceneDev& d = hs.dev;
    d.dmin[0] = p.domain.min.x;
    d.dmin[1] = p.domain.min.y;
    d.dmin[2] = p.domain.min.z;
    d.dmax[0] = p.domain.max.x;
    d.dmax[1] = p.domain.max.y;
    d.dmax[2] = p.domain.max.z;
    d.friction = p.domain.friction;
    d.g[0] = p.gravity.x;
    d.g[1] = p.gravity.y;
    d.g[2] = p.gravity.z;
    d.c0 = p.speed_of_sound;
    d.rho0 = p.rest_density;
    d.alpha = p.viscosity_constant;
    d.sigma = p.surface_tension_constant;
    d.vmax = p.max_velocity;
    d.h0 = p.base_smoothing_radius;
    d.r0 = p.base_particle_radius;
    d.m0 = p.base_particle_mass;
    d.gamma = p.tait_gamma;
    d.nuLam = p.laminar_viscosity;
    d.B = ((p.rest_density * p.speed_of_sound) * p.speed_of_sound) / p.tait_gamma;
    d.sloshA = p.slosh_amplitude;
    d.sloshT = p.slosh_period;
    d.dt = dt;
    d.halfDt = dt / 2.0F;
    d.gx = g.gx;
    d.gy = g.gy;
    d.gz = g.gz;
    d.cs[0] = g.cs[0];
    d.cs[1] = g.cs[1];
    d.cs[2] = g.cs[2];
    d.cellBase = static_cast<uint32_t>(cBase);
    d.nCells = static_cast<uint32_t>(cells);
    d.pOff = static_cast<uint32_t>(pOff);
    d.pCap = capacity;
    d.bOff = static_cast<uint32_t>(bOff);
    d.bCount = nb;
    d.flags = kSceneActive | (adaptive ? kSceneAdaptive : 0U) |
              ((p.flags & ASPH_FLAG_FIX_CELL_START) != 0U ? kSceneFixCellStart : 0U) |
              ((p.flags & ASPH_FLAG_KERNEL_MERGED_H) != 0U ? kSceneKernelMergedH : 0U) |
              ((p.flags & ASPH_FLAG_TOLERANCE_MATH) != 0U ? kSceneToleranceMath : 0U) |
              // refinement is the only writer of h and m; the uniform loops
              // divide by h0 through its reciprocal, valid for h0 in
              // [2^-40, 2^39]
              (adaptive || !sharedDivisor(p.base_smoothing_radius) ? 0U : kSceneUniformHM);
    d.tpb = p.threads_per_block;
    if (r != nullptr)
    {
        d.criterion = r->criterion_type;
        d.initialCooldown = r->initial_cooldown;
        d.cooldown = r->cooldown == 0U ? 1U : r->cooldown;
        d.maxCount = r->max_particle_count;
        d.minMassRatio = r->min_mass_ratio;
        d.maxMassRatio = r->max_mass_ratio;
        d.batchRatio = r->max_batch_ratio;
        d.eps = r->splitting.epsilon;
        d.alphaSplit = r->splitting.alpha;
        d.centerRatio = r->splitting.center_mass_ratio;
        d.vertexRatio = r->splitting.vertex_mass_ratio;
        d.velSplit = r->velocity.split;
        d.velMerge = r->velocity.merge;
        d.vortSplit = r->vorticity.split;
        d.vortMerge = r->vorticity.merge;
        d.ifSplit = r->interface_ratio.split;
        d.ifMerge = r->interface_ratio.merge;
        d.zoneC[0] = r->zone.center.x;
        d.zoneC[1] = r->zone.center.y;
        d.zoneC[2] = r->zone.center.z;
        d.zoneR = r->zone.radius;
        d.zonePathR = r->zone.path_radius;
        d.zonePathT = r->zone.path_period;
    }
    // SphSimulation ctor (SphSimulation.cu:36-74): v = 0, acc = g over the
    // capacity, radius = r0, h = h0, m = m0 for the initial particles.
    {
        std::vector<F4> pos(capacity, F4 {0.F, 0.F, 0.F, 0.F});
        for (uint32_t i = 0; i < count; ++i)
        {
            pos[i] = F4 {positions[i].x, positions[i].y, positions[i].z, positions[i].w};
        }
        std::vector<F4> zero4(capacity, F4 {0.F, 0.F, 0.F, 0.F});
        std::vector<F4> acc(capacity, F4 {p.gravity.x, p.gravity.y, p.gravity.z, 0.F});
        std::vector<float> zero(capacity, 0.F);
        std::vector<float> rad(capacity, 0.F);
        std::vector<float> hh(capacity, 0.F);
        std::vector<float> mm(capacity, 0.F);
        for (uint32_t i = 0; i < count; ++i)
        {
            rad[i] = p.base_particle_radius;
            hh[i] = p.base_smoothing_radius;
            mm[i] = p.base_particle_mass;
        }
        const size_t o = pOff;
        hipCheck(hipMemcpy(ctx_.pos + o, pos.data(), sizeof(F4) * capacity, hipMemcpyHostToDevice), "upload");
        hipCheck(hipMemcpy(ctx_.vel + o, zero4.data(), sizeof(F4) * capacity, hipMemcpyHostToDevice), "upload");
        hipCheck(hipMemcpy(ctx_.acc + o, acc.data(), sizeof(F4) * capacity, hipMemcpyHostToDevice), "upload");
        hipCheck(hipMemcpy(ctx_.rho + o, zero.data(), sizeof(float) * capacity, hipMemcpyHostToDevice), "upload");
        hipCheck(hipMemcpy(ctx_.rad + o, rad.data(), sizeof(float) * capacity, hipMemcpyHostToDevice), "upload");
        hipCheck(hipMemcpy(ctx_.h + o, hh.data(), sizeof(float) * capacity, hipMemcpyHostToDevice), "upload");
        hipCheck(hipMemcpy(ctx_.m + o, mm.data(), sizeof(float) * capacity, hipMemcpyHostToDevice), "upload");
    }
    hipCheck(hipMemset(ctx_.cellOff + cBase, 0, sizeof(uint32_t) * (cells + 1)), "hipMemset");
    uploadBoundary(id, boundary, nb);
    SceneState st {};
    st.count = count;
    hipCheck(hipMemcpy(ctx_.states + id, &st, sizeof(SceneState), hipMemcpyHostToDevice), "upload");
    uploadScene(id);
    uploadOrder();
    hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    return id;
}

void Engine::uploadBoundary(uint32_t id, const asph_boundary_particle* b, uint32_t nb)
{
    // NeighborGrid::updateBoundary (NeighborGrid.cu:42-49) done once on the
    // host: keys, stable sort, cell ranges with the slot-0 omission.
    HostScene& hs = scenes_[id];
    SceneDev& d = hs.dev;
    std::vector<int32_t> keys(nb);
    for (uint32_t i = 0; i < nb; ++i)
    {
        keys[i] = hostCellKey(d, b[i].position.x, b[i].position.y, b[i].position.z);
    }
    std::vector<uint32_t> order(nb);
    std::iota(order.begin(), order.end(), 0U);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t c) { return keys[a] < keys[c]; });
    std::vector<F4> s0(nb);
    std::vector<float> visc(nb, d.friction);
    hs.boundaryKeysSorted.resize(nb);
    for (uint32_t k = 0; k < nb; ++k)
    {
        const auto& q = b[order[k]];
        s0[k] = F4 {q.position.x, q.position.y, q.position.z, q.psi};
        hs.boundaryKeysSorted[k] = keys[order[k]];
    }
    hs.boundarySorted = s0;
    const uint32_t cells = d.nCells;
    std::vector<uint32_t> off(cells + 1, 0);
    for (uint32_t k = 0; k < nb; ++k)
    {
        off[static_cast<uint32_t>(hs.boundaryKeysSorted[k]) + 1U]++;
    }
    for (uint32_t c = 0; c < cells; ++c)
    {
        off[c + 1] += off[c];
    }
    for (auto& v : off)
    {
        v += d.bOff;
    }
    uint32_t lost = d.bOff;
    if ((d.flags & kSceneFixCellStart) == 0U && nb > 0)
    {
        if (nb == 1)
        {
            lost = d.bOff + 1;
        }
        else if (hs.boundaryKeysSorted[0] != hs.boundaryKeysSorted[1])
        {
            lost = off[static_cast<uint32_t>(hs.boundaryKeysSorted[1]) + 1U];
        }
    }
    d.bLostEnd = lost;
    // Boundary walks are skipped for queries whose cell box avoids the
    // outermost cells, which is exact only if every boundary particle lies in
    // an outermost cell (true for StaticBoundaryDomain walls outside the box).
    bool inside = false;
    for (uint32_t k = 0; k < nb; ++k)
    {
        const int32_t key = hs.boundaryKeysSorted[k];
        const int cx = key % d.gx;
        const int cy = (key / d.gx) % d.gy;
        const int cz = key / (d.gx * d.gy);
        if (cx != 0 && cy != 0 && cz != 0 && cx != d.gx - 1 && cy != d.gy - 1 && cz != d.gz - 1)
        {
            inside = true;
            break;
        }
    }
    d.flags = inside ? (d.flags | kSceneBoundaryInside) : (d.flags & ~kSceneBoundaryInside);
    d.bCount = nb;
    if (nb > 0)
    {
        hipCheck(hipMemcpy(ctx_.bs0 + d.bOff, s0.data(), sizeof(F4) * nb, hipMemcpyHostToDevice), "upload");
        hipCheck(hipMemcpy(ctx_.bvisc + d.bOff, visc.data(), sizeof(float) * nb, hipMemcpyHostToDevice), "upload");
    }
    hipCheck(hipMemcpy(ctx_.bcellOff + d.cellBase, off.data(), sizeof(uint32_t) * (cells + 1), hipMemcpyHostToDevice),
             "upload");
    // per-cell boxes of the boundary particles (all of the cell's particles,
    // so the box also covers slots the slot-0 omission hides)
    const float inf = std::numeric_limits<float>::infinity();
    std::vector<F4> lo(cells + 1, F4 {inf, inf, inf, 0.0F});
    std::vector<F4> hi(cells + 1, F4 {-inf, -inf, -inf, 0.0F});
    for (uint32_t k = 0; k < nb; ++k)
    {
        const uint32_t cell = static_cast<uint32_t>(hs.boundaryKeysSorted[k]);
        const F4& p = s0[k];
        lo[cell].x = std::min(lo[cell].x, p.x);
        lo[cell].y = std::min(lo[cell].y, p.y);
        lo[cell].z = std::min(lo[cell].z, p.z);
        hi[cell].x = std::max(hi[cell].x, p.x);
        hi[cell].y = std::max(hi[cell].y, p.y);
        hi[cell].z = std::max(hi[cell].z, p.z);
    }
    hipCheck(hipMemcpy(ctx_.bbox0 + d.cellBase, lo.data(), sizeof(F4) * (cells + 1), hipMemcpyHostToDevice), "upload");
    hipCheck(hipMemcpy(ctx_.bbox1 + d.cellBase, hi.data(), sizeof(F4) * (cells + 1), hipMemcpyHostToDevice), "upload");
    // Chunk boxes: each cell's boundary slots in chunks of kBoundaryGroup from
    // the cell's first slot, the box of a chunk stored at its first slot. A
    // cell at an edge or a corner of the domain collects several walls, so its
    // own box reaches into the fluid, while a chunk (consecutive slots of one
    // cell, in generation order) is a strip of one wall.
    if (nb > 0)
    {
        std::vector<F4> clo(nb, F4 {inf, inf, inf, 0.0F});
        std::vector<F4> chi(nb, F4 {-inf, -inf, -inf, 0.0F});
        std::vector<uint32_t> starts;
        for (uint32_t k = 0; k < nb; ++k)
        {
            const bool first = k == 0 || hs.boundaryKeysSorted[k] != hs.boundaryKeysSorted[k - 1];
            if (first || (k - starts.back()) == kBoundaryGroup)
            {
                starts.push_back(k);
            }
            const uint32_t k0 = starts.back();
            const F4& p = s0[k];
            clo[k0].x = std::min(clo[k0].x, p.x);
            clo[k0].y = std::min(clo[k0].y, p.y);
            clo[k0].z = std::min(clo[k0].z, p.z);
            chi[k0].x = std::max(chi[k0].x, p.x);
            chi[k0].y = std::max(chi[k0].y, p.y);
            chi[k0].z = std::max(chi[k0].z, p.z);
        }
        hipCheck(hipMemcpy(ctx_.bgbox0 + d.bOff, clo.data(), sizeof(F4) * nb, hipMemcpyHostToDevice), "upload");
        hipCheck(hipMemcpy(ctx_.bgbox1 + d.bOff, chi.data(), sizeof(F4) * nb, hipMemcpyHostToDevice), "upload");
        std::vector<F4> glo;
        std::vector<F4> ghi;
        glo.reserve(starts.size());
        ghi.reserve(starts.size());
        for (const uint32_t k0 : starts)
        {
            glo.push_back(clo[k0]);
            ghi.push_back(chi[k0]);
        }
        nearBounds(d, glo, ghi);
    }
    else
    {
        std::vector<float> far(cells, std::numeric_limits<float>::max());
        hipCheck(hipMemcpy(ctx_.bnear2 + d.cellBase, far.data(), sizeof(float) * cells, hipMemcpyHostToDevice),
                 "upload");
    }
}

// Per cell of the scene, a lower bound of the squared distance from any point
// of the cell (widened by a margin that covers the rounding of cell keys) to
// any boundary particle, from the boxes of the boundary chunks; capped at two
// cell widths. A particle whose squared support (2h)^2 lies below its cell's
// bound has no boundary particle within 2h, so its boundary walks add
// nothing and are skipped. The bound is rounded down with a relative margin
// far above the rounding of the walks' own distance tests.
void Engine::nearBounds(const SceneDev& d, const std::vector<F4>& glo, const std::vector<F4>& ghi)
{
    const uint32_t cells = d.nCells;
    const double cs[3] = {d.cs[0], d.cs[1], d.cs[2]};
    const double mn[3] = {d.dmin[0], d.dmin[1], d.dmin[2]};
    const int gdim[3] = {d.gx, d.gy, d.gz};
    const double cap = 2.0 * std::max(cs[0], std::max(cs[1], cs[2]));
    std::vector<double> near(cells, cap * cap);
    for (size_t g = 0; g < glo.size(); ++g)
    {
        const double lo[3] = {glo[g].x, glo[g].y, glo[g].z};
        const double hi[3] = {ghi[g].x, ghi[g].y, ghi[g].z};
        int i0[3];
        int i1[3];
        for (int a = 0; a < 3; ++a)
        {
            i0[a] = std::max(static_cast<int>(std::floor((lo[a] - cap - mn[a]) / cs[a])) - 1, 0);
            i1[a] = std::min(static_cast<int>(std::floor((hi[a] + cap - mn[a]) / cs[a])) + 1, gdim[a] - 1);
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
                        const double gap = std::max(0.0, std::max(lo[a] - ch, cl - hi[a]));
                        d2 += gap * gap;
                    }
                    const size_t cell =
                        static_cast<size_t>(x) + static_cast<size_t>(y) * gdim[0] +
                        static_cast<size_t>(z) * static_cast<size_t>(gdim[0]) * static_cast<size_t>(gdim[1]);
                    near[cell] = std::min(near[cell], d2);
                }
            }
        }
    }
    std::vector<float> out(cells);
    for (uint32_t k = 0; k < cells; ++k)
    {
        const float f = static_cast<float>(near[k] * (1.0 - 1e-5));
        out[k] = f > 0.0F ? std::nextafter(f, 0.0F) : 0.0F;
    }
    hipCheck(hipMemcpy(ctx_.bnear2 + d.cellBase, out.data(), sizeof(float) * cells, hipMemcpyHostToDevice), "upload");
}

void Engine::uploadScene(uint32_t id)
{
    hipCheck(hipMemcpyAsync(const_cast<SceneDev*>(ctx_.scenes) + id, &scenes_[id].dev, sizeof(SceneDev),
                            hipMemcpyHostToDevice, stream_),
             "upload scene");
    hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Engine::uploadOrder()
{
    std::vector<uint32_t> order;
    for (uint32_t s = 0; s < scenes_.size(); ++s)
    {
        if (scenes_[s].active)
        {
            order.push_back(s);
        }
    }
    std::sort(order.begin(), order.end(),
              [&](uint32_t a, uint32_t b) { return scenes_[a].dev.cellBase < scenes_[b].dev.cellBase; });
    const uint32_t n = static_cast<uint32_t>(order.size());
    if (n > 0)
    {
        std::vector<uint32_t> bases(n);
        for (uint32_t k = 0; k < n; ++k)
        {
            bases[k] = scenes_[order[k]].dev.cellBase;
        }
        hipCheck(hipMemcpy(cellOrderDev_, order.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice), "upload");
        hipCheck(hipMemcpy(cellBaseOrderDev_, bases.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice), "upload");
    }
    hipCheck(hipMemcpy(nOrderDev_, &n, sizeof(uint32_t), hipMemcpyHostToDevice), "upload");
    updateExtents();
}

void Engine::updateExtents()
{
    uint64_t nMax = 0;
    uint64_t cExtent = 2;
    uint32_t kinds = 0;
    for (const HostScene& hs : scenes_)
    {
        if (hs.active)
        {
            nMax += hs.dev.pCap;
            cExtent = std::max<uint64_t>(cExtent, static_cast<uint64_t>(hs.dev.cellBase) + hs.cellRegion);
            kinds |= toleranceForce(hs.dev) ? kForceTolerance : kForceExact;
        }
    }
    nMax = std::max<uint64_t>(nMax, 1);
    kinds = kinds == 0U ? kForceExact : kinds;
    const int bits = bitsFor(cExtent - 1);
    if (nMax == ctx_.nMax && cExtent == ctx_.cExtent && bits == keyBits_ && kinds == ctx_.forceKinds)
    {
        return;
    }
    // kernels take the context by value, so captured steps are re-captured
    ctx_.nMax = static_cast<uint32_t>(nMax);
    ctx_.cExtent = static_cast<uint32_t>(cExtent);
    ctx_.forceKinds = kinds;
    keyBits_ = bits;
    buildCtx();
    invalidateGraphs();
}

void Engine::removeScene(uint32_t id)
{
    if (id >= scenes_.size() || !scenes_[id].active)
    {
        throw Error(ASPH_ERR_INVALID, "no such scene");
    }
    synchronize();
    HostScene& hs = scenes_[id];
    pAlloc_.release(hs.dev.pOff, hs.dev.pCap);
    bAlloc_.release(hs.dev.bOff, hs.dev.bCount);
    cAlloc_.release(hs.dev.cellBase, hs.cellRegion);
    hs.active = false;
    hs.dev.flags = 0;
    SceneState st {};
    hipCheck(hipMemcpy(ctx_.states + id, &st, sizeof(SceneState), hipMemcpyHostToDevice), "upload");
    uploadScene(id);
    uploadOrder();
}

bool Engine::sceneActive(uint32_t id) const { return id < scenes_.size() && scenes_[id].active; }

uint32_t Engine::activeScenes() const
{
    uint32_t n = 0;
    for (const auto& s : scenes_)
    {
        n += s.active ? 1U : 0U;
    }
    return n;
}

void Engine::setDt(uint32_t id, float dt)
{
    HostScene& hs = scenes_.at(id);
    if (hs.dev.dt == dt)
    {
        return;
    }
    synchronize();
    hs.dev.dt = dt;
    hs.dev.halfDt = dt / 2.0F;  // SphSimulation.cu:103
    uploadScene(id);
}

void Engine::updateDomain(uint32_t id, const asph_domain& dom, const asph_boundary_particle* boundary, uint32_t nb)
{
    prepareDomain(id, dom, nb);
    uploadBoundary(id, boundary, nb);
    uploadScene(id);
    uploadOrder();
    hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

void Engine::updateDomainWalls(uint32_t id, const asph_domain& dom, float spacing, F4* latticePos, float* latticePsi)
{
    if (id >= scenes_.size() || !scenes_[id].active)
    {
        throw Error(ASPH_ERR_INVALID, "no such scene");
    }
    if (!(spacing > 0.0F))
    {
        throw Error(ASPH_ERR_INVALID, "particle spacing must be positive");
    }
    HostScene& hs0 = scenes_.at(id);
    (void)makeGrid(dom, hs0.dev.h0);  // refuse a grid beyond the limits before building walls
    const WallFaces wf = wallFaces(dom, spacing, hs0.dev.h0);
    prepareDomain(id, dom, wf.count);
    HostScene& hs = scenes_.at(id);
    if (wf.count == 0)
    {
        uploadBoundary(id, nullptr, 0);
    }
    else
    {
        SceneDev& d = hs.dev;
        d.bCount = wf.count;
        const WallTargets out {ctx_.bs0,    ctx_.bvisc,  ctx_.bcellOff, ctx_.bbox0,     ctx_.bbox1,
                               ctx_.bgbox0, ctx_.bgbox1, ctx_.bnear2,   latticePos, latticePsi};
        uint32_t lost = d.bOff;
        bool inside = false;
        wallsOnDevice(wf, d, out, &lost, &inside);
        d.bLostEnd = lost;
        d.flags = inside ? (d.flags | kSceneBoundaryInside) : (d.flags & ~kSceneBoundaryInside);
        hs.boundaryKeysSorted.clear();
        hs.boundarySorted.clear();
    }
    uploadScene(id);
    uploadOrder();
    hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

void Engine::prepareDomain(uint32_t id, const asph_domain& dom, uint32_t nb)
{
    HostScene& hs = scenes_.at(id);
    if (!(dom.max.x > dom.min.x) || !(dom.max.y > dom.min.y) || !(dom.max.z > dom.min.z))
    {
        throw Error(ASPH_ERR_INVALID, "domain.max must exceed domain.min on every axis");
    }
    // the same grid limits as at creation, before anything changes
    const GridGeom g = makeGrid(dom, hs.dev.h0);
    const uint64_t cells = g.cells();
    synchronize();
    // NeighborGrid::updateBoundarySize reallocates the cell arrays for the new
    // domain (NeighborGrid.cu:51-85). The new ranges are taken first fit with
    // the scene's old ranges counted as free. When one does not fit, the old
    // ranges are taken back (nothing else was allocated meanwhile) and each new
    // range is taken with its pool grown as needed (pool contents move, the
    // old ranges stay valid); the old ranges are released only once both new
    // ones are held, so a failure (ASPH_ERR_CAPACITY, a failed hipMalloc)
    // leaves the scene as it was.
    const uint64_t oldC = hs.dev.cellBase;
    const uint64_t oldCn = hs.cellRegion;
    const uint64_t oldB = hs.dev.bOff;
    const uint64_t oldBn = hs.dev.bCount;
    cAlloc_.release(oldC, oldCn);
    bAlloc_.release(oldB, oldBn);
    uint64_t cBase = cAlloc_.alloc(cells + 1);
    uint64_t bOff = bAlloc_.alloc(nb);
    if (cBase == UINT64_MAX || bOff == UINT64_MAX)
    {
        if (cBase != UINT64_MAX)
        {
            cAlloc_.release(cBase, cells + 1);
        }
        if (bOff != UINT64_MAX)
        {
            bAlloc_.release(bOff, nb);
        }
        if (!cAlloc_.claim(oldC, oldCn) || !bAlloc_.claim(oldB, oldBn))
        {
            throw Error(ASPH_ERR_STATE, "scene ranges lost while updating the domain");
        }
        cBase = cAlloc_.alloc(cells + 1);
        if (cBase == UINT64_MAX)
        {
            growCells(cAlloc_.capacity() + cells + 1);
            cBase = cAlloc_.alloc(cells + 1);
        }
        if (cBase == UINT64_MAX)
        {
            throw Error(ASPH_ERR_CAPACITY, "cell pool exhausted while updating the domain");
        }
        bOff = bAlloc_.alloc(nb);
        if (bOff == UINT64_MAX)
        {
            try
            {
                growBoundary(bAlloc_.capacity() + nb);
            }
            catch (...)
            {
                cAlloc_.release(cBase, cells + 1);
                throw;
            }
            bOff = bAlloc_.alloc(nb);
        }
        if (bOff == UINT64_MAX)
        {
            cAlloc_.release(cBase, cells + 1);
            throw Error(ASPH_ERR_CAPACITY, "boundary pool exhausted while updating the domain");
        }
        cAlloc_.release(oldC, oldCn);
        bAlloc_.release(oldB, oldBn);
    }
    SceneDev& d = hs.dev;
    d.dmin[0] = dom.min.x;
    d.dmin[1] = dom.min.y;
    d.dmin[2] = dom.min.z;
    d.dmax[0] = dom.max.x;
    d.dmax[1] = dom.max.y;
    d.dmax[2] = dom.max.z;
    d.friction = dom.friction;
    d.gx = g.gx;
    d.gy = g.gy;
    d.gz = g.gz;
    d.cs[0] = g.cs[0];
    d.cs[1] = g.cs[1];
    d.cs[2] = g.cs[2];
    d.cellBase = static_cast<uint32_t>(cBase);
    d.nCells = static_cast<uint32_t>(cells);
    d.bOff = static_cast<uint32_t>(bOff);
    hs.cellRegion = static_cast<uint32_t>(cells + 1);
    hipCheck(hipMemset(ctx_.cellOff + cBase, 0, sizeof(uint32_t) * (cells + 1)), "hipMemset");
}

template <class T>
void Engine::regrow(T*& p, size_t oldN, size_t newN)
{
    T* q = nullptr;
    hipCheck(hipMalloc(&q, std::max<size_t>(newN, 1) * sizeof(T)), "hipMalloc");
    hipCheck(hipMemset(q, 0, std::max<size_t>(newN, 1) * sizeof(T)), "hipMemset");
    if (oldN > 0)
    {
        hipCheck(hipMemcpy(q, p, oldN * sizeof(T), hipMemcpyDeviceToDevice), "hipMemcpy");
    }
    for (auto& a : allocations_)
    {
        if (a == p)
        {
            a = q;
        }
    }
    (void)hipFree(p);
    bytes_ += (newN - oldN) * sizeof(T);
    p = q;
}

void Engine::growCells(uint64_t newCap)
{
    synchronize();
    const size_t oldC = ctx_.cCap;
    const size_t C = std::max<size_t>(newCap, oldC);
    regrow(ctx_.cellOff, oldC + 1, C + 1);
    regrow(ctx_.bcellOff, oldC + 1, C + 1);
    regrow(ctx_.bbox0, oldC + 1, C + 1);
    regrow(ctx_.bbox1, oldC + 1, C + 1);
    regrow(ctx_.bnear2, oldC + 1, C + 1);
    regrow(tileSums_, scanTileWords(std::max<uint32_t>(ctx_.pCap, static_cast<uint32_t>(oldC) + 1)),
           scanTileWords(std::max<uint32_t>(ctx_.pCap, static_cast<uint32_t>(C) + 1)));
    ctx_.cCap = static_cast<uint32_t>(C);
    cAlloc_.grow(C);
    buildCtx();
    invalidateGraphs();
}

void Engine::growBoundary(uint64_t newCap)
{
    synchronize();
    const size_t oldB = std::max<uint64_t>(bAlloc_.capacity(), 1);
    const size_t B = std::max<size_t>(newCap, oldB);
    regrow(ctx_.bs0, oldB, B);
    regrow(ctx_.bvisc, oldB, B);
    regrow(ctx_.bgbox0, oldB, B);
    regrow(ctx_.bgbox1, oldB, B);
    bAlloc_.grow(B);
    invalidateGraphs();
}

SceneState Engine::readState(uint32_t id)
{
    synchronize();
    SceneState st {};
    hipCheck(hipMemcpy(&st, ctx_.states + id, sizeof(SceneState), hipMemcpyDeviceToHost), "download");
    return st;
}

void Engine::enableRefinement(uint32_t id)
{
    HostScene& hs = scenes_.at(id);
    if (!hs.adaptive)
    {
        return;  // SphSimulation::enableAdaptiveRefinement is a no-op (SphSimulation.cu:365)
    }
    setFrame(id, hs.dev.initialCooldown);  // AdaptiveSphSimulation.cu:87-90
}

void Engine::setFrame(uint32_t id, uint32_t frame)
{
    HostScene& hs = scenes_.at(id);
    SceneState st = readState(id);
    st.frame = frame;
    hs.frame = frame;
    hipCheck(hipMemcpy(ctx_.states + id, &st, sizeof(SceneState), hipMemcpyHostToDevice), "upload");
}

bool Engine::anyRefineNext() const
{
    for (const auto& s : scenes_)
    {
        if (s.active && s.adaptive)
        {
            const uint32_t f = s.frame;
            const uint32_t ic = s.dev.initialCooldown;
            if (f == ic || (f > ic && f % s.dev.cooldown == 0U))
            {
                return true;
            }
        }
    }
    return false;
}

void Engine::advanceFrames()
{
    for (auto& s : scenes_)
    {
        if (s.active && s.adaptive)
        {
            s.frame++;
        }
    }
}

const char* const kStageNames[] = {"prep",           "kick_drift_key", "cell_sort",    "gather",
                                   "neighbors",      "density",        "force",        "merge_criteria",
                                   "criterion_sort", "merge",          "compaction",   "split",
                                   "kick_refined"};
constexpr uint32_t kStageCount = sizeof(kStageNames) / sizeof(kStageNames[0]);

void Engine::enqueueStep(bool refine, std::vector<hipEvent_t>* marks)
{
    const Ctx& c = ctx_;
    const LaunchCfg& lc = cfg_.launch;
    hipStream_t s = stream_;
    auto mark = [&]() {
        if (marks != nullptr)
        {
            hipEvent_t e = nullptr;
            hipCheck(hipEventCreate(&e), "hipEventCreate");
            hipCheck(hipEventRecord(e, s), "hipEventRecord");
            marks->push_back(e);
        }
    };
    mark();
    launchPrep(c, s);
    mark();
    launchKickDriftKey(c, lc, s);
    mark();
    radixSortU32(c.keysA, c.valsA, c.keysB, c.valsB, c.nTotal, c.nMax, keyBits_,
                 radixDigitBits(lc.radixBits, keyBits_), hist_, digitTotals_,
                 s);
    mark();
    launchGather(c, lc, s);
    if (c.qperm != nullptr)
    {
        launchQueryOrder(c, s);
    }
    mark();
    launchNeighbors(c, lc, s);
    mark();
    launchDensity(c, lc, s);
    mark();
    launchForce(c, lc, s);
    mark();
    if (refine)
    {
        launchMergeCriteria(c, lc, s);
        mark();
        radixSortU64(c.ckeysA, c.cvalsA, c.ckeysB, c.cvalsB, c.critCount, c.nMax, 32 + sceneBits_, hist_, digitTotals_,
                     s);
        mark();
        launchMergeSelect(c, s);
        if (debugCapture_)
        {
            launchCopyMergeList(c, s);
        }
        launchMergeCandidates(c, lc, s);
        launchResolvePairs(c, lc, s);
        launchMergeApply(c, lc, s);
        mark();
        exclusiveScanU32(c.removal, c.rprefix, c.nTotal, c.nMax, tileSums_, scanTotal_, false, s);
        launchCompact(c, lc, s);
        launchAfterMerge(c, s);
        launchCopyBack(c, lc, s);
        mark();
        launchSplitCriteria(c, lc, s);
        radixSortU64(c.ckeysA, c.cvalsA, c.ckeysB, c.cvalsB, c.critCount, c.nMax, 32 + sceneBits_, hist_, digitTotals_,
                     s);
        launchSplitSelect(c, s);
        launchSplitApply(c, lc, s);
        mark();
        launchKickRefined(c, lc, s);
        mark();
    }
    hipCheck(hipGetLastError(), "kernel launch");
}

std::vector<float> Engine::profileStep()
{
    synchronize();
    const bool refine = anyRefineNext();
    std::vector<hipEvent_t> marks;
    enqueueStep(refine, &marks);
    advanceFrames();
    synchronize();
    std::vector<float> ms;
    for (size_t k = 0; k + 1 < marks.size(); ++k)
    {
        float t = 0.F;
        hipCheck(hipEventElapsedTime(&t, marks[k], marks[k + 1]), "hipEventElapsedTime");
        ms.push_back(t);
    }
    for (hipEvent_t e : marks)
    {
        (void)hipEventDestroy(e);
    }
    return ms;
}

void Engine::profileStepAsync()
{
    const bool refine = anyRefineNext();
    std::vector<hipEvent_t> marks;
    enqueueStep(refine, &marks);
    advanceFrames();
    pendingMarks_.push_back(std::move(marks));
}

std::vector<float> Engine::collectProfiles(uint32_t* samples)
{
    synchronize();
    std::vector<double> sum;
    std::vector<uint32_t> n;
    for (const auto& marks : pendingMarks_)
    {
        for (size_t k = 0; k + 1 < marks.size(); ++k)
        {
            float t = 0.F;
            hipCheck(hipEventElapsedTime(&t, marks[k], marks[k + 1]), "hipEventElapsedTime");
            if (sum.size() <= k)
            {
                sum.resize(k + 1, 0.0);
                n.resize(k + 1, 0U);
            }
            sum[k] += t;
            n[k] += 1U;
        }
        for (hipEvent_t e : marks)
        {
            (void)hipEventDestroy(e);
        }
    }
    *samples = static_cast<uint32_t>(pendingMarks_.size());
    pendingMarks_.clear();
    std::vector<float> ms(sum.size());
    for (size_t k = 0; k < sum.size(); ++k)
    {
        ms[k] = static_cast<float>(sum[k] / n[k]);
    }
    return ms;
}

float Engine::timedStep()
{
    hipEvent_t a = nullptr;
    hipEvent_t b = nullptr;
    hipCheck(hipEventCreate(&a), "hipEventCreate");
    hipCheck(hipEventCreate(&b), "hipEventCreate");
    hipCheck(hipEventRecord(a, stream_), "hipEventRecord");
    step(1);
    hipCheck(hipEventRecord(b, stream_), "hipEventRecord");
    hipCheck(hipEventSynchronize(b), "hipEventSynchronize");
    float t = 0.F;
    hipCheck(hipEventElapsedTime(&t, a, b), "hipEventElapsedTime");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return t;
}

const char* stageName(uint32_t i) { return i < kStageCount ? kStageNames[i] : ""; }

void Engine::step(uint32_t steps)
{
    for (uint32_t k = 0; k < steps; ++k)
    {
        const bool refine = anyRefineNext();
        if (cfg_.useGraph)
        {
            hipGraphExec_t& exec = refine ? graphRefine_ : graphPlain_;
            hipGraph_t& graph = refine ? graphRefineG_ : graphPlainG_;
            if (exec == nullptr)
            {
                hipCheck(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
                try
                {
                    enqueueStep(refine, nullptr);
                }
                catch (...)
                {
                    // leave the stream out of capture mode before reporting
                    hipGraph_t partial = nullptr;
                    (void)hipStreamEndCapture(stream_, &partial);
                    if (partial != nullptr)
                    {
                        (void)hipGraphDestroy(partial);
                    }
                    throw;
                }
                hipCheck(hipStreamEndCapture(stream_, &graph), "hipStreamEndCapture");
                hipCheck(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
            }
            hipCheck(hipGraphLaunch(exec, stream_), "hipGraphLaunch");
        }
        else
        {
            enqueueStep(refine, nullptr);
        }
        advanceFrames();
    }
}

void Engine::synchronize() { hipCheck(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

uint32_t Engine::count(uint32_t id) { return readState(id).count; }

uint32_t Engine::lastStepParticles()
{
    synchronize();
    uint32_t n = 0;
    hipCheck(hipMemcpy(&n, ctx_.nTotal, sizeof(n), hipMemcpyDeviceToHost), "download");
    return n;
}

uint32_t Engine::boundaryCount(uint32_t id) const { return scenes_.at(id).dev.bCount; }

uint32_t Engine::capacity(uint32_t id) const { return scenes_.at(id).dev.pCap; }

GridGeom Engine::gridGeom(uint32_t id) const
{
    const SceneDev& d = scenes_.at(id).dev;
    GridGeom g;
    g.gx = d.gx;
    g.gy = d.gy;
    g.gz = d.gz;
    g.cs[0] = d.cs[0];
    g.cs[1] = d.cs[1];
    g.cs[2] = d.cs[2];
    return g;
}

void Engine::copyFluid(uint32_t id, uint32_t mask, asph_vec4* pos, asph_vec4* vel, asph_vec4* acc, float* rho,
                       float* rad, float* h, float* m)
{
    const uint32_t n = count(id);
    const size_t o = scenes_.at(id).dev.pOff;
    auto get = [&](void* dst, const void* src, size_t bytes) {
        if (dst != nullptr && bytes > 0)
        {
            hipCheck(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "download");
        }
    };
    if ((mask & ASPH_FIELD_POSITIONS) != 0U) get(pos, ctx_.pos + o, sizeof(F4) * n);
    if ((mask & ASPH_FIELD_VELOCITIES) != 0U) get(vel, ctx_.vel + o, sizeof(F4) * n);
    if ((mask & ASPH_FIELD_ACCELERATIONS) != 0U) get(acc, ctx_.acc + o, sizeof(F4) * n);
    if ((mask & ASPH_FIELD_DENSITIES) != 0U) get(rho, ctx_.rho + o, sizeof(float) * n);
    if ((mask & ASPH_FIELD_RADII) != 0U) get(rad, ctx_.rad + o, sizeof(float) * n);
    if ((mask & ASPH_FIELD_SMOOTHING_RADII) != 0U) get(h, ctx_.h + o, sizeof(float) * n);
    if ((mask & ASPH_FIELD_MASSES) != 0U) get(m, ctx_.m + o, sizeof(float) * n);
}

void Engine::setFluid(uint32_t id, uint32_t count, const asph_vec4* pos, const asph_vec4* vel, const asph_vec4* acc,
                      const float* rho, const float* rad, const float* h, const float* m)
{
    HostScene& hs = scenes_.at(id);
    if (count > hs.dev.pCap)
    {
        throw Error(ASPH_ERR_CAPACITY, "count exceeds the scene capacity");
    }
    auto checkW = [&](const asph_vec4* a) {
        if (a == nullptr) return;
        for (uint32_t i = 0; i < count; ++i)
        {
            if (a[i].w != 0.0F)
            {
                throw Error(ASPH_ERR_INVALID, "vector fields must have w == 0");
            }
        }
    };
    checkW(pos);
    checkW(vel);
    checkW(acc);
    SceneState st = readState(id);
    const size_t o = hs.dev.pOff;
    auto put = [&](void* dst, const void* src, size_t bytes) {
        if (src != nullptr && bytes > 0)
        {
            hipCheck(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "upload");
        }
    };
    put(ctx_.pos + o, pos, sizeof(F4) * count);
    put(ctx_.vel + o, vel, sizeof(F4) * count);
    put(ctx_.acc + o, acc, sizeof(F4) * count);
    put(ctx_.rho + o, rho, sizeof(float) * count);
    put(ctx_.rad + o, rad, sizeof(float) * count);
    put(ctx_.h + o, h, sizeof(float) * count);
    put(ctx_.m + o, m, sizeof(float) * count);
    st.count = count;
    hipCheck(hipMemcpy(ctx_.states + id, &st, sizeof(SceneState), hipMemcpyHostToDevice), "upload");
    // a plain scene keeps the uniform-h/m fast path while every particle has
    // exactly h0 and m0
    auto known = [&](const float* a, float v, uint32_t prev) {
        if (a == nullptr)
        {
            return prev;
        }
        for (uint32_t i = 0; i < count; ++i)
        {
            if (std::memcmp(&a[i], &v, sizeof(float)) != 0)
            {
                return i;
            }
        }
        return std::max(prev, count);
    };
    hs.knownH = known(h, hs.dev.h0, hs.knownH);
    hs.knownM = known(m, hs.dev.m0, hs.knownM);
    const bool uniform = !hs.adaptive && hs.knownH >= count && hs.knownM >= count && sharedDivisor(hs.dev.h0);
    const uint32_t flags = uniform ? (hs.dev.flags | kSceneUniformHM) : (hs.dev.flags & ~kSceneUniformHM);
    if (flags != hs.dev.flags)
    {
        hs.dev.flags = flags;
        uploadScene(id);
    }
}

asph_fluid_view Engine::view(uint32_t id)
{
    const size_t o = scenes_.at(id).dev.pOff;
    asph_fluid_view v {};
    v.positions = reinterpret_cast<asph_vec4*>(ctx_.pos + o);
    v.velocities = reinterpret_cast<asph_vec4*>(ctx_.vel + o);
    v.accelerations = reinterpret_cast<asph_vec4*>(ctx_.acc + o);
    v.densities = ctx_.rho + o;
    v.radii = ctx_.rad + o;
    v.smoothing_radii = ctx_.h + o;
    v.masses = ctx_.m + o;
    v.particle_count = count(id);
    return v;
}

float Engine::averageNeighborCount(uint32_t id)
{
    // calculateAverageNeighborCount (SphSimulation.cu:237-247)
    const uint32_t n = count(id);
    hipCheck(hipMemsetAsync(diagTotal_, 0, sizeof(uint32_t), stream_), "hipMemsetAsync");
    if (n > 0)
    {
        launchNeighborSets(ctx_, id, n, nullptr, nullptr, diagTotal_, stream_);
    }
    uint32_t total = 0;
    hipCheck(hipMemcpyAsync(&total, diagTotal_, sizeof(uint32_t), hipMemcpyDeviceToHost, stream_), "download");
    synchronize();
    return n > 0 ? static_cast<float>(total) / static_cast<float>(n) : 0.F;
}

void Engine::neighborSets(uint32_t id, uint32_t* counts, uint64_t* prints)
{
    const uint32_t n = count(id);
    if (n == 0)
    {
        return;
    }
    uint32_t* dc = nullptr;
    uint64_t* dp = nullptr;
    hipCheck(hipMalloc(&dc, sizeof(uint32_t) * n), "hipMalloc");
    hipCheck(hipMalloc(&dp, sizeof(uint64_t) * n), "hipMalloc");
    hipCheck(hipMemsetAsync(diagTotal_, 0, sizeof(uint32_t), stream_), "hipMemsetAsync");
    launchNeighborSets(ctx_, id, n, dc, dp, diagTotal_, stream_);
    synchronize();
    hipError_t e1 = hipSuccess;
    if (counts != nullptr) e1 = hipMemcpy(counts, dc, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
    hipError_t e2 = hipSuccess;
    if (prints != nullptr) e2 = hipMemcpy(prints, dp, sizeof(uint64_t) * n, hipMemcpyDeviceToHost);
    (void)hipFree(dc);
    (void)hipFree(dp);
    hipCheck(e1, "download");
    hipCheck(e2, "download");
}

asph_density_info Engine::densityInfo(uint32_t id, float threshold)
{
    // getDensityInfo (SphSimulation.cu:249-285); the sum is accumulated in
    // FP64 in a fixed order.
    const uint32_t n = count(id);
    const SceneDev& d = scenes_.at(id).dev;
    asph_density_info info {d.rho0, 3.402823466e+38F, -3.402823466e+38F, 0.F, 0U, 0U, 0U};
    launchDensityInfo(ctx_, id, n, threshold, infoD_, infoU_, stream_);
    std::vector<double> dp(kInfoBlocks);
    std::vector<uint32_t> up(5 * kInfoBlocks);
    hipCheck(hipMemcpyAsync(dp.data(), infoD_, sizeof(double) * kInfoBlocks, hipMemcpyDeviceToHost, stream_),
             "download");
    hipCheck(hipMemcpyAsync(up.data(), infoU_, sizeof(uint32_t) * 5 * kInfoBlocks, hipMemcpyDeviceToHost, stream_),
             "download");
    synchronize();
    double sum = 0.0;
    for (uint32_t b = 0; b < kInfoBlocks; ++b)
    {
        float mn = 0.F;
        float mx = 0.F;
        std::memcpy(&mn, &up[5 * b], sizeof(float));
        std::memcpy(&mx, &up[5 * b + 1], sizeof(float));
        info.min_density = std::min(info.min_density, mn);
        info.max_density = std::max(info.max_density, mx);
        info.under_density_count += up[5 * b + 2];
        info.normal_density_count += up[5 * b + 3];
        info.over_density_count += up[5 * b + 4];
        sum += dp[b];
    }
    info.average_density = static_cast<float>(sum / static_cast<double>(n));
    return info;
}

void Engine::grid(uint32_t id, int32_t* keys, int32_t* idx, int32_t* cellStart, int32_t* cellEnd)
{
    const SceneState st = readState(id);
    const SceneDev& d = scenes_.at(id).dev;
    const uint32_t n = st.count;
    const uint32_t ss = st.sortedStart;
    std::vector<uint32_t> k(n);
    std::vector<uint32_t> v(n);
    if (n > 0)
    {
        hipCheck(hipMemcpy(k.data(), ctx_.sortedKeys + ss, sizeof(uint32_t) * n, hipMemcpyDeviceToHost), "download");
        hipCheck(hipMemcpy(v.data(), ctx_.sortedVals + ss, sizeof(uint32_t) * n, hipMemcpyDeviceToHost), "download");
    }
    for (uint32_t i = 0; i < n; ++i)
    {
        if (keys != nullptr) keys[i] = static_cast<int32_t>(k[i] - d.cellBase);
        if (idx != nullptr) idx[i] = static_cast<int32_t>(v[i] - d.pOff);
    }
    if (cellStart == nullptr && cellEnd == nullptr)
    {
        return;
    }
    std::vector<uint32_t> off(d.nCells + 1);
    hipCheck(hipMemcpy(off.data(), ctx_.cellOff + d.cellBase, sizeof(uint32_t) * (d.nCells + 1),
                       hipMemcpyDeviceToHost),
             "download");
    // Reference representation (inclusive start/end, -1 when unset) of the
    // device cell ranges and the omitted prefix [ss, lostEnd).
    for (uint32_t c = 0; c < d.nCells; ++c)
    {
        const uint32_t a = off[c];
        const uint32_t b = off[c + 1];
        int32_t s = -1;
        int32_t e = -1;
        if (b > a)
        {
            const uint32_t lo = std::max(a, st.lostEnd);
            if (b > lo)
            {
                s = static_cast<int32_t>(lo - ss);
                e = static_cast<int32_t>(b - 1 - ss);
            }
            else if (n == 1)
            {
                s = -1;
                e = 0;
            }
            else if (a == ss)
            {
                s = 0;
                e = -1;
            }
            else
            {
                s = -1;
                e = static_cast<int32_t>(b - 1 - ss);
            }
        }
        if (cellStart != nullptr) cellStart[c] = s;
        if (cellEnd != nullptr) cellEnd[c] = e;
    }
}

void Engine::refinementTrace(uint32_t id, asph_refinement_trace* t, uint32_t* mergeList, uint32_t* cand,
                             uint32_t* pairKeep, uint32_t* splitParents)
{
    const SceneState st = readState(id);
    asph_refinement_trace out {};
    out.refined = st.refined;
    out.count_before = st.countBefore;
    out.merge_eligible = st.mergeK;
    out.merge_pairs = st.mergePairs;
    out.count_after_merge = st.countAfterMerge;
    out.split_selected = st.splitK;
    out.split_performed = st.splitDone;
    out.count_after = st.countAfter;
    if (t != nullptr)
    {
        *t = out;
    }
    if (!cfg_.adaptive || st.refined == 0U)
    {
        return;
    }
    if (mergeList != nullptr && st.mergeK > 0)
    {
        if (!debugCapture_)
        {
            throw Error(ASPH_ERR_STATE, "merge list capture is off");
        }
        hipCheck(hipMemcpy(mergeList, ctx_.mergeListDbg, sizeof(uint32_t) * st.mergeK, hipMemcpyDeviceToHost),
                 "download");
    }
    const bool merged = st.countBefore > 1U;
    if (cand != nullptr && merged)
    {
        hipCheck(hipMemcpy(cand, ctx_.cand + st.denseOff, sizeof(uint32_t) * st.countBefore, hipMemcpyDeviceToHost),
                 "download");
    }
    if (pairKeep != nullptr && merged && st.mergeK > 0)
    {
        std::vector<uint32_t> kp(st.countBefore);
        hipCheck(hipMemcpy(kp.data(), ctx_.keep + st.denseOff, sizeof(uint32_t) * st.countBefore,
                           hipMemcpyDeviceToHost),
                 "download");
        uint32_t w = 0;
        for (uint32_t i = 0; i < st.countBefore; ++i)
        {
            if (kp[i] != 0U)
            {
                pairKeep[w++] = i;
            }
        }
    }
    if (splitParents != nullptr && st.splitDone > 0)
    {
        hipCheck(hipMemcpy(splitParents, ctx_.csortedVals + st.segStart, sizeof(uint32_t) * st.splitDone,
                           hipMemcpyDeviceToHost),
                 "download");
    }
}

}  // namespace asph
