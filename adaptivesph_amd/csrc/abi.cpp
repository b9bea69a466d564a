// The C ABI (include/asph.h, asph_debug.h, asph_batch.h) over the engine.
// No exception crosses this boundary: each entry point returns a status and
// records the message for asph_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/asph.h"
#include "../../include/asph_batch.h"
#include "../../include/asph_debug.h"
#include "engine.hpp"

#include "walls.hpp"

struct asph_simulation
{
    std::unique_ptr<asph::Engine> engine;
    uint32_t scene = 0;
    bool adaptive = false;
    float lastDt = -1.0F;
    // BoundaryParticlesData in the boundary domain's order: positions,
    // colours and radii in the imported boundary buffers or in arrays owned
    // here; psi and viscosity coefficients owned here
    asph_vec4* bPos = nullptr;
    asph_vec4* bColors = nullptr;
    float* bRadii = nullptr;
    float* bPsi = nullptr;
    float* bVisc = nullptr;
    uint32_t bCap = 0;
    uint32_t bCount = 0;
    bool bImported = false;  // positions, colours and radii are the caller's (fixed capacity)
    float radius = 0.0F;
    std::vector<void*> owned;

    ~asph_simulation()
    {
        engine.reset();  // finishes the simulation's work before its buffers go
        for (void* q : owned)
        {
            (void)hipFree(q);
        }
    }
};

struct asph_batch
{
    std::unique_ptr<asph::Engine> engine;
};

namespace
{

thread_local std::string g_error;

template <class F>
int guard(F&& f)
{
    try
    {
        f();
        g_error.clear();
        return ASPH_OK;
    }
    catch (const asph::Error& e)
    {
        g_error = e.what();
        return e.code;
    }
    catch (const std::bad_alloc&)
    {
        g_error = "host allocation failed";
        return ASPH_ERR_CAPACITY;
    }
    catch (const std::exception& e)
    {
        g_error = e.what();
        return ASPH_ERR_INVALID;
    }
}

void need(const void* p, const char* what)
{
    if (p == nullptr)
    {
        throw asph::Error(ASPH_ERR_INVALID, std::string(what) + " is NULL");
    }
}

asph::Engine& engineOf(asph_simulation* s)
{
    need(s, "sim");
    return *s->engine;
}

asph::Engine& engineOf(asph_batch* b)
{
    need(b, "batch");
    return *b->engine;
}

void checkScene(asph::Engine& e, uint32_t id)
{
    if (!e.sceneActive(id))
    {
        throw asph::Error(ASPH_ERR_INVALID, "no such scene");
    }
}

void* ownedAlloc(asph_simulation* sim, size_t bytes)
{
    void* q = nullptr;
    asph::hipCheck(hipMalloc(&q, std::max<size_t>(bytes, 1)), "hipMalloc");
    sim->owned.push_back(q);
    return q;
}

// Replaces an owned array by a new one of `bytes` (contents not kept).
template <class T>
void ownedRealloc(asph_simulation* sim, T*& p, size_t bytes)
{
    T* q = static_cast<T*>(ownedAlloc(sim, bytes));
    for (auto it = sim->owned.begin(); it != sim->owned.end(); ++it)
    {
        if (*it == static_cast<void*>(p))
        {
            sim->owned.erase(it);
            (void)hipFree(p);
            break;
        }
    }
    p = q;
}

void needBuffer(const asph_device_buffer& b, size_t bytes, const char* what)
{
    if (b.data == nullptr || b.size < bytes)
    {
        throw asph::Error(ASPH_ERR_CAPACITY, std::string(what) + " is NULL or smaller than the capacity needs (" +
                                                 std::to_string(bytes) + " bytes)");
    }
}

// Boundary arrays of a new simulation (capacity cap): the imported buffers
// when given, else arrays owned by the simulation.
void setupBoundary(asph_simulation* sim, const asph_boundary_buffers* mem, uint32_t cap, float radius)
{
    const size_t n = std::max<uint32_t>(cap, 1);
    sim->bCap = cap;
    sim->radius = radius;
    if (mem != nullptr)
    {
        needBuffer(mem->positions, sizeof(asph_vec4) * cap, "boundary positions buffer");
        needBuffer(mem->radii, sizeof(float) * cap, "boundary radii buffer");
        needBuffer(mem->colors, sizeof(asph_vec4) * cap, "boundary colours buffer");
        sim->bPos = static_cast<asph_vec4*>(mem->positions.data);
        sim->bRadii = static_cast<float*>(mem->radii.data);
        sim->bColors = static_cast<asph_vec4*>(mem->colors.data);
        sim->bImported = true;
    }
    else
    {
        sim->bPos = static_cast<asph_vec4*>(ownedAlloc(sim, sizeof(asph_vec4) * n));
        sim->bRadii = static_cast<float*>(ownedAlloc(sim, sizeof(float) * n));
        sim->bColors = static_cast<asph_vec4*>(ownedAlloc(sim, sizeof(asph_vec4) * n));
    }
    sim->bPsi = static_cast<float*>(ownedAlloc(sim, sizeof(float) * n));
    sim->bVisc = static_cast<float*>(ownedAlloc(sim, sizeof(float) * n));
}

// Room for n wall particles: imported boundary buffers have a fixed
// capacity (ASPH_ERR_CAPACITY beyond it, before anything changes); the
// simulation's own arrays grow, as its boundary pools do.
void needBoundaryCapacity(asph_simulation* sim, uint64_t n)
{
    if (n <= sim->bCap)
    {
        return;
    }
    if (sim->bImported || n > UINT32_MAX)
    {
        throw asph::Error(ASPH_ERR_CAPACITY, "the domain has " + std::to_string(n) +
                                                 " boundary particles, beyond the boundary capacity " +
                                                 std::to_string(sim->bCap) + " of the imported buffers");
    }
    ownedRealloc(sim, sim->bPos, sizeof(asph_vec4) * n);
    ownedRealloc(sim, sim->bColors, sizeof(asph_vec4) * n);
    ownedRealloc(sim, sim->bRadii, sizeof(float) * n);
    ownedRealloc(sim, sim->bPsi, sizeof(float) * n);
    ownedRealloc(sim, sim->bVisc, sizeof(float) * n);
    sim->bCap = static_cast<uint32_t>(n);
}

// SphSimulation::initializeBoundaryParticles (SphSimulation.cu:287-333): the
// radii (the base particle radius), colours (0.3, 0.3, 0.3, 0.2) and
// viscosity coefficients (domain.friction) of n wall particles; positions and
// psi from the records when given (else already on the device).
void writeBoundary(asph_simulation* sim, const asph_boundary_particle* records, uint32_t n, float friction)
{
    sim->bCount = n;
    if (n == 0)
    {
        return;
    }
    if (records != nullptr)
    {
        std::vector<asph_vec4> pos(n);
        std::vector<float> psi(n);
        for (uint32_t i = 0; i < n; ++i)
        {
            pos[i] = records[i].position;
            psi[i] = records[i].psi;
        }
        asph::hipCheck(hipMemcpy(sim->bPos, pos.data(), sizeof(asph_vec4) * n, hipMemcpyHostToDevice), "upload");
        asph::hipCheck(hipMemcpy(sim->bPsi, psi.data(), sizeof(float) * n, hipMemcpyHostToDevice), "upload");
    }
    const std::vector<float> radii(n, sim->radius);
    const std::vector<asph_vec4> colors(n, asph_vec4 {0.3F, 0.3F, 0.3F, 0.2F});
    const std::vector<float> visc(n, friction);
    asph::hipCheck(hipMemcpy(sim->bRadii, radii.data(), sizeof(float) * n, hipMemcpyHostToDevice), "upload");
    asph::hipCheck(hipMemcpy(sim->bColors, colors.data(), sizeof(asph_vec4) * n, hipMemcpyHostToDevice), "upload");
    asph::hipCheck(hipMemcpy(sim->bVisc, visc.data(), sizeof(float) * n, hipMemcpyHostToDevice), "upload");
}

void createImpl(const asph_parameters* params, const asph_vec4* positions, uint32_t count,
                const asph_fluid_buffers* fluidMem, const asph_boundary_buffers* boundaryMem,
                const asph_boundary_particle* boundary, uint32_t boundary_count,
                const asph_refinement_parameters* refinement, uint32_t max_fluid_capacity,
                uint32_t max_boundary_capacity, asph_simulation** out)
{
    need(params, "params");
    need(out, "out");
    *out = nullptr;
    asph::validateParams(*params);
    const bool adaptive = refinement != nullptr && refinement->enabled != 0U;
    const uint32_t cap = adaptive ? refinement->max_particle_count : max_fluid_capacity;
    if (count > cap)
    {
        throw asph::Error(ASPH_ERR_CAPACITY, "particle count exceeds the fluid capacity");
    }
    if (cap > asph::kMaxParticleCapacity)
    {
        throw asph::Error(ASPH_ERR_CAPACITY, "particle capacity above 2^28 - 64");
    }
    if (boundary_count > max_boundary_capacity)
    {
        throw asph::Error(ASPH_ERR_CAPACITY, "boundary count exceeds the boundary capacity");
    }
    if (boundary_count > 0)
    {
        need(boundary, "boundary");
    }
    const asph::GridGeom g = asph::makeGrid(params->domain, params->base_smoothing_radius);
    asph::EngineConfig cfg;
    cfg.particleCapacity = cap;
    cfg.boundaryCapacity = max_boundary_capacity;
    cfg.cellCapacity = static_cast<uint32_t>(g.cells() + 1U);
    cfg.sceneSlots = 1;
    cfg.listCap = asph::defaultListCap(adaptive);
    if (adaptive)
    {
        // the list of a very large capacity may not take more than a
        // quarter of the free device memory (never below the plain
        // default; particles past the capacity take the walking path)
        size_t freeBytes = 0;
        size_t totalBytes = 0;
        if (hipMemGetInfo(&freeBytes, &totalBytes) == hipSuccess && cap > 0U)
        {
            const uint64_t fit = static_cast<uint64_t>(freeBytes) / 4U / (4ULL * cap);
            const uint64_t k = std::max<uint64_t>(asph::kDefaultListCapPlain, fit & ~3ULL);
            cfg.listCap = static_cast<uint32_t>(std::min<uint64_t>(cfg.listCap, k));
        }
    }
    if (fluidMem != nullptr)
    {
        // FluidParticlesDataImportedBuffer: the simulation's own arrays
        needBuffer(fluidMem->positions, sizeof(asph_vec4) * cap, "fluid positions buffer");
        needBuffer(fluidMem->velocities, sizeof(asph_vec4) * cap, "fluid velocities buffer");
        needBuffer(fluidMem->densities, sizeof(float) * cap, "fluid densities buffer");
        needBuffer(fluidMem->radii, sizeof(float) * cap, "fluid radii buffer");
        cfg.importedPos = static_cast<asph::F4*>(fluidMem->positions.data);
        cfg.importedVel = static_cast<asph::F4*>(fluidMem->velocities.data);
        cfg.importedRho = static_cast<float*>(fluidMem->densities.data);
        cfg.importedRad = static_cast<float*>(fluidMem->radii.data);
    }
    cfg.adaptive = adaptive;
    cfg.useGraph = false;
    auto sim = std::make_unique<asph_simulation>();
    setupBoundary(sim.get(), boundaryMem, max_boundary_capacity, params->base_particle_radius);
    sim->engine = std::make_unique<asph::Engine>(cfg);
    sim->engine->setDebugCapture(adaptive);
    sim->scene = sim->engine->addScene(*params, positions, count, boundary, boundary_count, refinement, cap, 0.0F);
    sim->adaptive = adaptive;
    writeBoundary(sim.get(), boundary, boundary_count, params->domain.friction);
    *out = sim.release();
}

}  // namespace

extern "C" {

const char* asph_last_error(void) { return g_error.c_str(); }

int asph_default_parameters(asph_parameters* p)
{
    return guard([&] {
        need(p, "p");
        std::memset(p, 0, sizeof(*p));
        p->domain.min = {-1.F, -1.F, -1.F};
        p->domain.max = {1.F, 1.F, 1.F};
        p->domain.friction = 0.01F;
        p->threads_per_block = 256;
        p->tait_gamma = 5.0F;
    });
}

int asph_default_refinement_parameters(asph_refinement_parameters* r)
{
    return guard([&] {
        need(r, "r");
        std::memset(r, 0, sizeof(*r));
        r->enabled = 0;
        r->min_mass_ratio = 0.3F;
        r->max_mass_ratio = 8.0F;
        r->max_particle_count = 1000000;
        r->max_batch_ratio = 0.2F;
        r->initial_cooldown = 10;
        r->cooldown = 10;
        r->criterion_type = ASPH_CRITERION_VELOCITY;
        r->splitting = {0.65F, 0.70F, 0.077F, 0.077F};
        r->velocity = {2.0F, 0.1F};
        r->vorticity = {1.0F, 0.3F};
        r->curvature = {2.0F, 0.6F};
        r->interface_ratio = {0.05F, 0.15F};
    });
}

int asph_boundary_generate(const asph_domain* domain, float particle_spacing, float rest_density,
                           float smoothing_radius, asph_boundary_particle* out, uint32_t* count)
{
    return guard([&] {
        need(domain, "domain");
        need(count, "count");
        if (!(particle_spacing > 0.0F) || !(smoothing_radius > 0.0F))
        {
            throw asph::Error(ASPH_ERR_INVALID, "spacing and smoothing radius must be positive");
        }
        const int rc =
            asph::boundaryGenerate(*domain, particle_spacing, rest_density, smoothing_radius, out, count, false);
        if (rc != ASPH_OK)
        {
            throw asph::Error(rc, "output array too small");
        }
    });
}

int asph_boundary_generate_device(const asph_domain* domain, float particle_spacing, float rest_density,
                                  float smoothing_radius, asph_boundary_particle* out, uint32_t* count)
{
    return guard([&] {
        need(domain, "domain");
        need(count, "count");
        if (!(particle_spacing > 0.0F) || !(smoothing_radius > 0.0F))
        {
            throw asph::Error(ASPH_ERR_INVALID, "spacing and smoothing radius must be positive");
        }
        const int rc =
            asph::boundaryGenerate(*domain, particle_spacing, rest_density, smoothing_radius, out, count, true);
        if (rc != ASPH_OK)
        {
            throw asph::Error(rc, "output array too small");
        }
    });
}

int asph_boundary_generate_device_noalloc(const asph_domain* domain, float particle_spacing, float rest_density,
                                          float smoothing_radius, asph_boundary_particle* out, uint32_t* count,
                                          void* stream)
{
    return guard([&] {
        need(domain, "domain");
        need(count, "count");
        if (!(particle_spacing > 0.0F) || !(smoothing_radius > 0.0F))
        {
            throw asph::Error(ASPH_ERR_INVALID, "spacing and smoothing radius must be positive");
        }
        const int rc = asph::boundaryGenerateNoAlloc(*domain, particle_spacing, rest_density, smoothing_radius, out,
                                                     count, static_cast<hipStream_t>(stream));
        if (rc != ASPH_OK)
        {
            throw asph::Error(rc, "output array too small");
        }
    });
}

int asph_create(const asph_parameters* params, const asph_vec4* positions, uint32_t count,
                const asph_boundary_particle* boundary, uint32_t boundary_count,
                const asph_refinement_parameters* refinement, uint32_t max_fluid_capacity,
                uint32_t max_boundary_capacity, asph_simulation** out)
{
    return guard([&] {
        createImpl(params, positions, count, nullptr, nullptr, boundary, boundary_count, refinement,
                   max_fluid_capacity, max_boundary_capacity, out);
    });
}

int asph_create_imported(const asph_parameters* params, const asph_vec4* positions, uint32_t count,
                         const asph_fluid_buffers* fluid_memory, const asph_boundary_buffers* boundary_memory,
                         const asph_boundary_particle* boundary, uint32_t boundary_count,
                         const asph_refinement_parameters* refinement, uint32_t max_fluid_capacity,
                         uint32_t max_boundary_capacity, asph_simulation** out)
{
    return guard([&] {
        need(fluid_memory, "fluid_memory");
        need(boundary_memory, "boundary_memory");
        createImpl(params, positions, count, fluid_memory, boundary_memory, boundary, boundary_count, refinement,
                   max_fluid_capacity, max_boundary_capacity, out);
    });
}

int asph_buffer_allocate(size_t size, asph_device_buffer* out)
{
    return guard([&] {
        need(out, "out");
        out->data = nullptr;
        out->size = 0;
        void* q = nullptr;
        asph::hipCheck(hipMalloc(&q, std::max<size_t>(size, 1)), "hipMalloc");
        out->data = q;
        out->size = size;
    });
}

int asph_buffer_release(asph_device_buffer* buffer)
{
    return guard([&] {
        need(buffer, "buffer");
        if (buffer->data != nullptr)
        {
            asph::hipCheck(hipFree(buffer->data), "hipFree");
        }
        buffer->data = nullptr;
        buffer->size = 0;
    });
}

int asph_buffer_download(const asph_device_buffer* buffer, size_t offset, void* host, size_t bytes)
{
    return guard([&] {
        need(buffer, "buffer");
        need(host, "host");
        if (offset > buffer->size || bytes > buffer->size - offset)
        {
            throw asph::Error(ASPH_ERR_INVALID, "copy beyond the buffer");
        }
        asph::hipCheck(hipMemcpy(host, static_cast<const char*>(buffer->data) + offset, bytes, hipMemcpyDeviceToHost),
                       "download");
    });
}

int asph_buffer_upload(const asph_device_buffer* buffer, size_t offset, const void* host, size_t bytes)
{
    return guard([&] {
        need(buffer, "buffer");
        need(host, "host");
        if (offset > buffer->size || bytes > buffer->size - offset)
        {
            throw asph::Error(ASPH_ERR_INVALID, "copy beyond the buffer");
        }
        asph::hipCheck(hipMemcpy(static_cast<char*>(buffer->data) + offset, host, bytes, hipMemcpyHostToDevice),
                       "upload");
    });
}

int asph_get_boundary_view(asph_simulation* sim, asph_boundary_view* out)
{
    return guard([&] {
        need(sim, "sim");
        need(out, "out");
        out->positions = sim->bPos;
        out->colors = sim->bColors;
        out->psi_values = sim->bPsi;
        out->viscosity_coefficients = sim->bVisc;
        out->radii = sim->bRadii;
        out->particle_count = sim->bCount;
    });
}

int asph_destroy(asph_simulation* sim)
{
    return guard([&] { delete sim; });
}

int asph_update(asph_simulation* sim, float dt)
{
    return guard([&] {
        asph::Engine& e = engineOf(sim);
        if (!std::isfinite(dt))
        {
            throw asph::Error(ASPH_ERR_INVALID, "dt must be finite");
        }
        if (dt != sim->lastDt)
        {
            e.setDt(sim->scene, dt);
            sim->lastDt = dt;
        }
        e.step(1);
        e.synchronize();  // update is synchronous (SphSimulation.cu:119)
    });
}

int asph_update_domain(asph_simulation* sim, const asph_domain* domain, const asph_boundary_particle* boundary,
                       uint32_t boundary_count)
{
    return guard([&] {
        need(domain, "domain");
        if (boundary_count > 0)
        {
            need(boundary, "boundary");
        }
        asph::Engine& e = engineOf(sim);
        needBoundaryCapacity(sim, boundary_count);
        e.updateDomain(sim->scene, *domain, boundary, boundary_count);
        writeBoundary(sim, boundary, boundary_count, domain->friction);
    });
}

int asph_update_domain_walls(asph_simulation* sim, const asph_domain* domain, float particle_spacing)
{
    return guard([&] {
        need(domain, "domain");
        asph::Engine& e = engineOf(sim);
        // the new wall count, checked against the boundary arrays first
        asph::WallSequences sq;
        if (!(particle_spacing > 0.0F) ||
            asph::wallSequences(*domain, particle_spacing, 1ULL << 31, &sq) != asph::kWallSeqOk)
        {
            throw asph::Error(ASPH_ERR_INVALID, "particle spacing must be positive and advance the wall lattice");
        }
        asph::WallFaces wf;
        asph::wallFaceTable(sq, 1.0F, &wf);
        needBoundaryCapacity(sim, wf.count);
        e.updateDomainWalls(sim->scene, *domain, particle_spacing, reinterpret_cast<asph::F4*>(sim->bPos), sim->bPsi);
        writeBoundary(sim, nullptr, wf.count, domain->friction);
    });
}

int asph_fluid_count(asph_simulation* sim, uint32_t* count)
{
    return guard([&] {
        need(count, "count");
        *count = engineOf(sim).count(sim->scene);
    });
}

int asph_boundary_count(asph_simulation* sim, uint32_t* count)
{
    return guard([&] {
        need(count, "count");
        *count = engineOf(sim).boundaryCount(sim->scene);
    });
}

int asph_average_neighbor_count(asph_simulation* sim, float* out)
{
    return guard([&] {
        need(out, "out");
        *out = engineOf(sim).averageNeighborCount(sim->scene);
    });
}

int asph_get_density_info(asph_simulation* sim, float threshold, asph_density_info* out)
{
    return guard([&] {
        need(out, "out");
        *out = engineOf(sim).densityInfo(sim->scene, threshold);
    });
}

int asph_enable_adaptive_refinement(asph_simulation* sim)
{
    return guard([&] { engineOf(sim).enableRefinement(sim->scene); });
}

int asph_get_fluid_view(asph_simulation* sim, asph_fluid_view* out)
{
    return guard([&] {
        need(out, "out");
        *out = engineOf(sim).view(sim->scene);
    });
}

int asph_copy_fluid(asph_simulation* sim, uint32_t mask, asph_vec4* positions, asph_vec4* velocities,
                    asph_vec4* accelerations, float* densities, float* radii, float* smoothing_radii, float* masses)
{
    return guard([&] {
        engineOf(sim).copyFluid(sim->scene, mask, positions, velocities, accelerations, densities, radii,
                                smoothing_radii, masses);
    });
}

// ------------------------------------------------------------------ debug --

int asph_debug_set_fluid(asph_simulation* sim, uint32_t count, const asph_vec4* positions,
                         const asph_vec4* velocities, const asph_vec4* accelerations, const float* densities,
                         const float* radii, const float* smoothing_radii, const float* masses)
{
    return guard([&] {
        engineOf(sim).setFluid(sim->scene, count, positions, velocities, accelerations, densities, radii,
                               smoothing_radii, masses);
    });
}

int asph_debug_set_frame(asph_simulation* sim, uint32_t frame)
{
    return guard([&] { engineOf(sim).setFrame(sim->scene, frame); });
}

int asph_debug_grid_dims(asph_simulation* sim, int32_t dims[3])
{
    return guard([&] {
        need(dims, "dims");
        const asph::GridGeom g = engineOf(sim).gridGeom(sim->scene);
        dims[0] = g.gx;
        dims[1] = g.gy;
        dims[2] = g.gz;
    });
}

int asph_debug_grid(asph_simulation* sim, int32_t* sorted_keys, int32_t* sorted_indices, int32_t* cell_start,
                    int32_t* cell_end)
{
    return guard([&] { engineOf(sim).grid(sim->scene, sorted_keys, sorted_indices, cell_start, cell_end); });
}

int asph_debug_neighbor_sets(asph_simulation* sim, uint32_t* counts, uint64_t* fingerprints)
{
    return guard([&] { engineOf(sim).neighborSets(sim->scene, counts, fingerprints); });
}

int asph_debug_refinement(asph_simulation* sim, asph_refinement_trace* trace, uint32_t* merge_list,
                          uint32_t* merge_candidates, uint32_t* pair_keep_ids, uint32_t* split_parents)
{
    return guard([&] {
        engineOf(sim).refinementTrace(sim->scene, trace, merge_list, merge_candidates, pair_keep_ids, split_parents);
    });
}

int asph_debug_division_check(uint64_t seed, uint64_t count, uint64_t* mismatches)
{
    return guard([&] {
        need(mismatches, "mismatches");
        unsigned long long* bad = nullptr;
        asph::hipCheck(hipMalloc(&bad, sizeof(*bad)), "division check");
        unsigned long long h = 0;
        hipError_t e = hipMemset(bad, 0, sizeof(*bad));
        if (e == hipSuccess)
        {
            asph::launchDivisionCheck(seed, count, bad, nullptr);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
        {
            e = hipMemcpy(&h, bad, sizeof(h), hipMemcpyDeviceToHost);
        }
        (void)hipFree(bad);
        asph::hipCheck(e, "division check");
        *mismatches = h;
    });
}

// ------------------------------------------------------------------ batch --

int asph_batch_create(const asph_batch_config* config, asph_batch** out)
{
    return guard([&] {
        need(config, "config");
        need(out, "out");
        *out = nullptr;
        asph::EngineConfig cfg;
        cfg.particleCapacity = config->max_particles;
        cfg.boundaryCapacity = config->max_boundary;
        cfg.cellCapacity = config->max_cells;
        cfg.sceneSlots = config->max_scenes == 0 ? 1U : config->max_scenes;
        cfg.adaptive = (config->flags & ASPH_BATCH_ADAPTIVE) != 0U;
        cfg.listCap = config->list_capacity == 0 ? asph::defaultListCap(cfg.adaptive) : config->list_capacity;
        cfg.useGraph = (config->flags & ASPH_BATCH_GRAPH) != 0U;
        auto b = std::make_unique<asph_batch>();
        b->engine = std::make_unique<asph::Engine>(cfg);
        *out = b.release();
    });
}

int asph_batch_destroy(asph_batch* batch)
{
    return guard([&] { delete batch; });
}

int asph_batch_add_scene(asph_batch* batch, const asph_parameters* params, const asph_vec4* positions,
                         uint32_t count, const asph_boundary_particle* boundary, uint32_t boundary_count,
                         const asph_refinement_parameters* refinement, uint32_t capacity, float dt,
                         uint32_t* scene_id)
{
    return guard([&] {
        need(params, "params");
        need(scene_id, "scene_id");
        if (!(dt > 0.0F) || !std::isfinite(dt))
        {
            throw asph::Error(ASPH_ERR_INVALID, "dt must be positive and finite");
        }
        *scene_id = engineOf(batch).addScene(*params, positions, count, boundary, boundary_count, refinement,
                                             capacity, dt);
    });
}

int asph_batch_remove_scene(asph_batch* batch, uint32_t scene_id)
{
    return guard([&] { engineOf(batch).removeScene(scene_id); });
}

int asph_batch_update_domain_walls(asph_batch* batch, uint32_t scene_id, const asph_domain* domain,
                                   float particle_spacing)
{
    return guard([&] {
        need(domain, "domain");
        engineOf(batch).updateDomainWalls(scene_id, *domain, particle_spacing);
    });
}

int asph_batch_step(asph_batch* batch, uint32_t steps)
{
    return guard([&] { engineOf(batch).step(steps); });
}

int asph_batch_synchronize(asph_batch* batch)
{
    return guard([&] { engineOf(batch).synchronize(); });
}

int asph_batch_step_timed(asph_batch* batch, float* milliseconds)
{
    return guard([&] {
        need(milliseconds, "milliseconds");
        *milliseconds = engineOf(batch).timedStep();
    });
}

int asph_batch_step_particles(asph_batch* batch, uint32_t* count)
{
    return guard([&] {
        need(count, "count");
        *count = engineOf(batch).lastStepParticles();
    });
}

int asph_batch_average_neighbor_count(asph_batch* batch, uint32_t scene_id, float* out)
{
    return guard([&] {
        need(out, "out");
        asph::Engine& e = engineOf(batch);
        checkScene(e, scene_id);
        *out = e.averageNeighborCount(scene_id);
    });
}

int asph_batch_scene_count(asph_batch* batch, uint32_t scene_id, uint32_t* count)
{
    return guard([&] {
        need(count, "count");
        asph::Engine& e = engineOf(batch);
        checkScene(e, scene_id);
        *count = e.count(scene_id);
    });
}

int asph_batch_copy_scene(asph_batch* batch, uint32_t scene_id, uint32_t mask, asph_vec4* positions,
                          asph_vec4* velocities, asph_vec4* accelerations, float* densities, float* radii,
                          float* smoothing_radii, float* masses)
{
    return guard([&] {
        asph::Engine& e = engineOf(batch);
        checkScene(e, scene_id);
        e.copyFluid(scene_id, mask, positions, velocities, accelerations, densities, radii, smoothing_radii, masses);
    });
}

int asph_batch_set_schedule(asph_batch* batch, int32_t query_tile, uint32_t force_blocks_per_cu)
{
    return guard([&] {
        asph::Engine& e = engineOf(batch);
        if (query_tile < -1 || query_tile > 4)
        {
            throw asph::Error(ASPH_ERR_INVALID, "query_tile must be -1 (default), 0, 1, 2, 3 or 4");
        }
        if (force_blocks_per_cu > 64U)
        {
            throw asph::Error(ASPH_ERR_INVALID, "force_blocks_per_cu must be at most 64");
        }
        asph::LaunchCfg lc = e.config().launch;
        lc.queryTile = query_tile;
        lc.forceBlocksPerCU = force_blocks_per_cu;
        e.setLaunch(lc);
    });
}

int asph_batch_set_launch(asph_batch* batch, uint32_t block_stream, uint32_t block_neighbor, uint32_t radix_bits,
                          uint32_t use_graph)
{
    return guard([&] {
        asph::Engine& e = engineOf(batch);
        if (block_stream < 64 || block_stream > 1024 || (block_stream % 64) != 0)
        {
            throw asph::Error(ASPH_ERR_INVALID, "block_stream must be a multiple of 64 in [64, 1024]");
        }
        if (block_neighbor != 64 && block_neighbor != 128 && block_neighbor != 256)
        {
            throw asph::Error(ASPH_ERR_INVALID, "block_neighbor must be 64, 128 or 256");
        }
        if (radix_bits != 0 && radix_bits != 4 && radix_bits != 6 && radix_bits != 7 && radix_bits != 8 &&
            radix_bits != 10)
        {
            throw asph::Error(ASPH_ERR_INVALID, "radix_bits must be 0 (from the key width), 4, 6, 7, 8 or 10");
        }
        asph::LaunchCfg lc = e.config().launch;
        lc.blockStream = block_stream;
        lc.blockNeighbor = block_neighbor;
        lc.radixBits = static_cast<int>(radix_bits);
        e.setLaunch(lc);
        e.setUseGraph(use_graph != 0U);
    });
}

int asph_batch_memory(asph_batch* batch, uint64_t* allocated, uint64_t* particles_used)
{
    return guard([&] {
        asph::Engine& e = engineOf(batch);
        if (allocated != nullptr)
        {
            *allocated = e.bytesAllocated();
        }
        if (particles_used != nullptr)
        {
            *particles_used = e.particlesUsed();
        }
    });
}

int asph_batch_profile_step(asph_batch* batch, float* ms, uint32_t capacity, uint32_t* count)
{
    return guard([&] {
        need(count, "count");
        const std::vector<float> t = engineOf(batch).profileStep();
        *count = static_cast<uint32_t>(t.size());
        for (uint32_t i = 0; i < t.size() && i < capacity && ms != nullptr; ++i)
        {
            ms[i] = t[i];
        }
    });
}

int asph_batch_profile_step_async(asph_batch* batch)
{
    return guard([&] { engineOf(batch).profileStepAsync(); });
}

int asph_batch_profile_collect(asph_batch* batch, float* ms, uint32_t capacity, uint32_t* count, uint32_t* samples)
{
    return guard([&] {
        need(count, "count");
        need(samples, "samples");
        const std::vector<float> t = engineOf(batch).collectProfiles(samples);
        *count = static_cast<uint32_t>(t.size());
        for (uint32_t i = 0; i < t.size() && i < capacity && ms != nullptr; ++i)
        {
            ms[i] = t[i];
        }
    });
}

const char* asph_profile_stage_name(uint32_t stage) { return asph::stageName(stage); }

uint64_t asph_scene_bytes(uint32_t capacity, uint32_t boundary_count, uint32_t cells, uint32_t adaptive,
                          uint32_t list_capacity)
{
    return asph::sceneBytes(capacity, boundary_count, cells, adaptive != 0U,
                            list_capacity == 0 ? asph::defaultListCap(adaptive != 0U) : list_capacity);
}

int asph_grid_cells(const asph_domain* domain, float smoothing_radius, uint32_t* cells)
{
    return guard([&] {
        need(domain, "domain");
        need(cells, "cells");
        *cells = static_cast<uint32_t>(asph::makeGrid(*domain, smoothing_radius).cells());
    });
}

int asph_device_info(char* buffer, uint32_t length)
{
    return guard([&] {
        need(buffer, "buffer");
        int dev = 0;
        asph::hipCheck(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t prop {};
        asph::hipCheck(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        std::snprintf(buffer, length, "%s (%s), %d CUs, %.1f GiB", prop.name, prop.gcnArchName,
                      prop.multiProcessorCount, static_cast<double>(prop.totalGlobalMem) / (1024.0 * 1024.0 * 1024.0));
    });
}

}  // extern "C"
