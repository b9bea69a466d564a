// Host side of the HIP engine: device pools, the scene table, and the
// per-step launch sequence (optionally captured once as a hipGraph and
// replayed). A Simulation of the reference interface is an Engine holding one
// scene; a batch is an Engine holding many.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/asph.h"
#include "../../include/asph_debug.h"
#include "kernels.hpp"

namespace asph
{

struct Error : std::runtime_error
{
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hipCheck(hipError_t e, const char* what);

// First-fit allocator of index ranges inside a fixed pool.
class RangeAllocator
{
public:
    explicit RangeAllocator(uint64_t capacity = 0);
    // Returns the offset of a free range of n elements, or UINT64_MAX.
    uint64_t alloc(uint64_t n);
    // Takes the free range [off, off + n) (false if any of it is in use).
    bool claim(uint64_t off, uint64_t n);
    void release(uint64_t off, uint64_t n);
    void grow(uint64_t capacity);  // appends [capacity(), capacity) as free
    uint64_t capacity() const { return cap_; }
    uint64_t used() const { return used_; }

private:
    struct Range
    {
        uint64_t off, n;
    };
    std::vector<Range> free_;
    uint64_t cap_ = 0;
    uint64_t used_ = 0;
};

// Grid geometry of NeighborGrid::createGrid (NeighborGrid.cu:108-121).
struct GridGeom
{
    int gx = 0, gy = 0, gz = 0;
    float cs[3] = {0.F, 0.F, 0.F};
    uint64_t cells() const { return static_cast<uint64_t>(gx) * static_cast<uint64_t>(gy) * static_cast<uint64_t>(gz); }
};
// Largest neighbour grid a scene may have (cells).
constexpr uint64_t kMaxGridCells = 1000000000ULL;
// Throws Error(ASPH_ERR_INVALID) when the domain is empty on an axis or the
// grid would have more than kMaxGridCells cells.
GridGeom makeGrid(const asph_domain& d, float h0);
// Throws Error(ASPH_ERR_INVALID) for parameters the pipeline cannot run.
void validateParams(const asph_parameters& p);

// Default neighbour-list entries per particle: plain scenes hold about 80
// neighbours (h = 1.3 dx); refinement packs up to 13 daughters per parent
// volume, and refined regions hold several hundred neighbours per particle.
// Particles beyond the capacity take the slower walking overflow kernels.
// Largest particle pool: record byte offsets (j * 16) must fit in 32 bits.
constexpr uint32_t kMaxParticleCapacity = ASPH_MAX_PARTICLE_CAPACITY;
static_assert(kMaxParticleCapacity == (1U << 28) - 64U, "record offsets j * 16 fit 32 bits");
constexpr uint32_t kDefaultListCapPlain = 128;
constexpr uint32_t kDefaultListCapAdaptive = 640;

inline uint32_t defaultListCap(bool adaptive) { return adaptive ? kDefaultListCapAdaptive : kDefaultListCapPlain; }

struct EngineConfig
{
    uint32_t particleCapacity = 0;  // pool of fluid particle slots
    uint32_t boundaryCapacity = 0;  // pool of boundary particle slots
    uint32_t cellCapacity = 0;      // pool of grid cells
    uint32_t sceneSlots = 1;        // scene table size
    uint32_t listCap = 128;         // neighbour-list entries per particle (rounded up to 4s)
    bool adaptive = false;          // allocate refinement scratch
    bool useGraph = false;          // capture and replay one step as a hipGraph
    LaunchCfg launch;
    // Caller-owned canonical arrays (imported particle memory, the buffers the
    // reference's simulation writes for the renderer), each of particleCapacity
    // records; null: allocated by the engine. The engine never frees them.
    F4* importedPos = nullptr;
    F4* importedVel = nullptr;
    float* importedRho = nullptr;
    float* importedRad = nullptr;
};

// Bytes of device memory a scene of this shape needs inside an engine
// (used for admission against a memory budget).
const char* stageName(uint32_t i);

uint64_t sceneBytes(uint32_t capacity, uint32_t boundaryCount, uint32_t cells, bool adaptive, uint32_t listCap);
// Neighbour-list capacity actually used: a whole number of 4-entry chunks.
uint32_t listCapacityRounded(uint32_t listCap);

class Engine
{
public:
    explicit Engine(const EngineConfig& cfg);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Adds a scene; returns its slot id. capacity is the fluid capacity.
    uint32_t addScene(const asph_parameters& p, const asph_vec4* positions, uint32_t count,
                      const asph_boundary_particle* boundary, uint32_t nb, const asph_refinement_parameters* r,
                      uint32_t capacity, float dt);
    void removeScene(uint32_t id);
    bool sceneActive(uint32_t id) const;
    void setDt(uint32_t id, float dt);
    void updateDomain(uint32_t id, const asph_domain& d, const asph_boundary_particle* boundary, uint32_t nb);
    // updateDomain with the walls of StaticBoundaryDomain::generate(d,
    // spacing, rho0, h0) built on the device (no particle data through the host)
    // latticePos / latticePsi (optional, device, at least the new wall count):
    // the walls' positions and psi in the generator's order.
    void updateDomainWalls(uint32_t id, const asph_domain& d, float spacing, F4* latticePos = nullptr,
                           float* latticePsi = nullptr);
    void enableRefinement(uint32_t id);
    void setFrame(uint32_t id, uint32_t frame);

    // Enqueue `steps` steps of every active scene (asynchronous).
    void step(uint32_t steps);
    void synchronize();

    uint32_t count(uint32_t id);  // synchronous
    // Particles of all scenes at the start of the most recent step (synchronous).
    uint32_t lastStepParticles();
    uint32_t boundaryCount(uint32_t id) const;
    uint32_t capacity(uint32_t id) const;
    void copyFluid(uint32_t id, uint32_t mask, asph_vec4* pos, asph_vec4* vel, asph_vec4* acc, float* rho,
                   float* rad, float* h, float* m);
    void setFluid(uint32_t id, uint32_t count, const asph_vec4* pos, const asph_vec4* vel, const asph_vec4* acc,
                  const float* rho, const float* rad, const float* h, const float* m);
    asph_fluid_view view(uint32_t id);
    float averageNeighborCount(uint32_t id);
    asph_density_info densityInfo(uint32_t id, float threshold);
    void neighborSets(uint32_t id, uint32_t* counts, uint64_t* prints);
    void grid(uint32_t id, int32_t* keys, int32_t* idx, int32_t* cellStart, int32_t* cellEnd);
    GridGeom gridGeom(uint32_t id) const;
    void refinementTrace(uint32_t id, asph_refinement_trace* t, uint32_t* mergeList, uint32_t* cand,
                         uint32_t* pairKeep, uint32_t* splitParents);
    void setDebugCapture(bool on) { debugCapture_ = on; }

    uint64_t bytesAllocated() const { return bytes_; }
    uint64_t particlesUsed() const { return pAlloc_.used(); }
    uint32_t activeScenes() const;
    const EngineConfig& config() const { return cfg_; }
    hipStream_t stream() const { return stream_; }
    void setLaunch(const LaunchCfg& lc);
    // One eager step with an event between stages (stage times in ms).
    std::vector<float> profileStep();
    // The same step left in flight (no wait); collectProfiles() waits for the
    // stream and returns each stage's time averaged over the steps so
    // enqueued since the last collection (their number in *samples).
    void profileStepAsync();
    std::vector<float> collectProfiles(uint32_t* samples);
    // One step bracketed by events (ms).
    float timedStep();
    void setUseGraph(bool on);

private:
    struct HostScene
    {
        bool active = false;
        SceneDev dev {};
        uint32_t cellRegion = 0;  // nCells + 1
        bool adaptive = false;
        uint32_t frame = 0;  // host mirror of the device frame counter
        uint32_t count0 = 0;
        // leading slots known to hold h = h0 (knownH) and m = m0 (knownM)
        uint32_t knownH = 0;
        uint32_t knownM = 0;
        std::vector<F4> boundarySorted;
        std::vector<int32_t> boundaryKeysSorted;
    };

    template <class T>
    T* alloc(size_t n);
    void buildCtx();
    void uploadScene(uint32_t id);
    void uploadOrder();
    void updateExtents();
    void uploadBoundary(uint32_t id, const asph_boundary_particle* b, uint32_t nb);
    // the new domain's grid and pool regions for nb boundary particles
    void prepareDomain(uint32_t id, const asph_domain& d, uint32_t nb);
    void nearBounds(const SceneDev& d, const std::vector<F4>& glo, const std::vector<F4>& ghi);
    void enqueueStep(bool refine, std::vector<hipEvent_t>* marks);
    bool anyRefineNext() const;
    void advanceFrames();
    void invalidateGraphs();
    template <class T>
    void regrow(T*& p, size_t oldN, size_t newN);
    void growCells(uint64_t newCap);
    void growBoundary(uint64_t newCap);
    SceneState readState(uint32_t id);

    EngineConfig cfg_;
    hipStream_t stream_ = nullptr;
    std::vector<void*> allocations_;
    uint64_t bytes_ = 0;
    Ctx ctx_ {};
    int keyBits_ = 1;
    int sortParity_ = 0;
    int csortParity_ = 0;
    int sceneBits_ = 1;
    RangeAllocator pAlloc_, bAlloc_, cAlloc_;
    std::vector<HostScene> scenes_;
    // device scratch not in Ctx
    uint32_t* hist_ = nullptr;
    uint32_t* digitTotals_ = nullptr;
    uint32_t* tileSums_ = nullptr;
    uint32_t* scanTotal_ = nullptr;
    uint32_t* nOrderDev_ = nullptr;
    uint32_t* qpermBuf_ = nullptr;  // query order buffers (Ctx::qperm / qpos when a query tile is set)
    uint32_t* qposBuf_ = nullptr;
    uint32_t computeUnits_ = 1;
    uint32_t* cellOrderDev_ = nullptr;
    uint32_t* cellBaseOrderDev_ = nullptr;
    double* infoD_ = nullptr;
    uint32_t* infoU_ = nullptr;
    uint32_t* diagTotal_ = nullptr;
    bool debugCapture_ = false;
    std::vector<std::vector<hipEvent_t>> pendingMarks_;  // profileStepAsync's events
    hipGraphExec_t graphPlain_ = nullptr;
    hipGraphExec_t graphRefine_ = nullptr;
    hipGraph_t graphPlainG_ = nullptr;
    hipGraph_t graphRefineG_ = nullptr;
};

}  // namespace asph
