// C++ interface of the simulation, header-only over the C ABI (asph.h).
//
// Mirrors the reference's public C++ surface so that a caller of
//   cuda/include/cuda/simulation/Simulation.cuh          (Simulation, Parameters,
//                                                         DensityInfo, createSimulation)
//   cuda/include/cuda/simulation/physics/StaticBoundaryDomain.cuh
//   cuda/include/cuda/refinement/RefinementParameters.cuh
// keeps its names: namespace sph::cuda, the same classes, methods, fields and
// defaults. Differences, all at the edges of that surface:
//   * glm::vec3 / glm::vec4 are sph::cuda::vec3 / vec4 (same layout) unless
//     glm is available, in which case they are glm's own types;
//   * ImportedMemory (ImportedMemory.cuh:10-26) is the same abstract base,
//     but importBuffer(fd, size) (Vulkan external memory) is replaced by
//     allocateBuffer(size), a device allocation, and wrapBuffer(ptr, size)
//     for device memory the caller already has; createSimulation takes the
//     reference's FluidParticlesDataImportedBuffer and
//     BoundaryParticlesDataImportedBuffer of such buffers (the 8-argument
//     form), or none (the simulation then owns its memory);
//   * fluidParticles() / boundaryParticles() return the device pointers in
//     FluidParticlesData / BoundaryParticlesData layout;
//   * errors arrive as std::runtime_error carrying the library's message, as
//     the reference's host failures arrive as exceptions.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../asph.h"

#if defined(__has_include)
#if __has_include(<glm/ext/vector_float3.hpp>) && __has_include(<glm/ext/vector_float4.hpp>)
#include <glm/ext/vector_float3.hpp>
#include <glm/ext/vector_float4.hpp>
#define SPH_HAVE_GLM 1
#endif
#endif

namespace sph::cuda
{

#ifdef SPH_HAVE_GLM
using vec3 = glm::vec3;
using vec4 = glm::vec4;
#else
struct vec3
{
    float x = 0.F, y = 0.F, z = 0.F;
};
struct vec4
{
    float x = 0.F, y = 0.F, z = 0.F, w = 0.F;
};
#endif
static_assert(sizeof(vec3) == sizeof(asph_vec3), "vec3 must be 3 floats");
static_assert(sizeof(vec4) == sizeof(asph_vec4), "vec4 must be 4 floats");

namespace detail
{
inline void check(int status, const char* what)
{
    if (status != ASPH_OK)
    {
        throw std::runtime_error(std::string(what) + ": " + asph_last_error());
    }
}
inline asph_vec3 toC(const vec3& v) { return {v.x, v.y, v.z}; }
}  // namespace detail

// ImportedMemory (ImportedMemory.cuh:10-26): device memory the caller hands
// the simulation. devicePointer() and size() describe it; copyToHost and
// copyFromHost move bytes.
class ImportedMemory
{
public:
    ImportedMemory() = default;
    ImportedMemory(const ImportedMemory&) = delete;
    ImportedMemory(ImportedMemory&&) = delete;
    ImportedMemory& operator=(const ImportedMemory&) = delete;
    ImportedMemory& operator=(ImportedMemory&&) = delete;
    virtual ~ImportedMemory() = default;

    [[nodiscard]] virtual void* devicePointer() const = 0;
    [[nodiscard]] virtual size_t size() const = 0;

    void copyToHost(void* host, size_t bytes, size_t offset = 0) const
    {
        const asph_device_buffer b = buffer();
        detail::check(asph_buffer_download(&b, offset, host, bytes), "ImportedMemory::copyToHost");
    }
    void copyFromHost(const void* host, size_t bytes, size_t offset = 0) const
    {
        const asph_device_buffer b = buffer();
        detail::check(asph_buffer_upload(&b, offset, host, bytes), "ImportedMemory::copyFromHost");
    }
    [[nodiscard]] asph_device_buffer buffer() const { return {devicePointer(), size()}; }
};

namespace detail
{
// A device allocation owned by the object (allocateBuffer).
class DeviceMemory final : public ImportedMemory
{
public:
    explicit DeviceMemory(size_t bytes) { check(asph_buffer_allocate(bytes, &b_), "allocateBuffer"); }
    ~DeviceMemory() override { (void)asph_buffer_release(&b_); }
    [[nodiscard]] void* devicePointer() const override { return b_.data; }
    [[nodiscard]] size_t size() const override { return b_.size; }

private:
    asph_device_buffer b_ {};
};

// Device memory owned by someone else (wrapBuffer).
class BorrowedMemory final : public ImportedMemory
{
public:
    BorrowedMemory(void* p, size_t bytes) : p_(p), n_(bytes) {}
    [[nodiscard]] void* devicePointer() const override { return p_; }
    [[nodiscard]] size_t size() const override { return n_; }

private:
    void* p_;
    size_t n_;
};
}  // namespace detail

// importBuffer (ImportedMemory.cuh:22-26) for a headless program: a device
// allocation of `size` bytes on the current GPU.
inline std::unique_ptr<ImportedMemory> allocateBuffer(size_t size)
{
    return std::make_unique<detail::DeviceMemory>(size);
}

// Device memory of the current GPU that the caller owns (it must outlive
// every simulation it is handed to).
inline std::unique_ptr<ImportedMemory> wrapBuffer(void* devicePointer, size_t size)
{
    return std::make_unique<detail::BorrowedMemory>(devicePointer, size);
}

// FluidParticlesDataImportedBuffer / BoundaryParticlesDataImportedBuffer
// (Simulation.cuh:46-59).
struct FluidParticlesDataImportedBuffer
{
    const ImportedMemory& positions;
    const ImportedMemory& velocities;
    const ImportedMemory& densities;
    const ImportedMemory& radii;
};

struct BoundaryParticlesDataImportedBuffer
{
    const ImportedMemory& positions;
    const ImportedMemory& radii;
    const ImportedMemory& colors;
};

// BoundaryParticlesData (Simulation.cuh:35-44): device pointers, in the
// boundary domain's order.
struct BoundaryParticlesData
{
    vec4* positions;
    vec4* colors;
    float* psiValues;
    float* viscosityCoefficients;
    float* radii;

    uint32_t particleCount;
};

// FluidParticlesData (Simulation.cuh:22-33): device pointers.
struct FluidParticlesData
{
    vec4* positions;
    vec4* velocities;
    vec4* accelerations;
    float* densities;
    float* radii;
    float* smoothingRadii;
    float* masses;

    uint32_t particleCount;
};

namespace refinement
{

// RefinementParameters.cuh:8-104, with the same defaults.
struct VelocityParameters
{
    struct Split
    {
        float minimalSpeedThreshold = 2.0F;
    };
    struct Merge
    {
        float maximalSpeedThreshold = 0.1F;
    };
    Split split;
    Merge merge;
};

struct VorticityParameters
{
    struct Split
    {
        float minimalVorticityThreshold = 1.0F;
    };
    struct Merge
    {
        float maximalVorticityThreshold = 0.3F;
    };
    Split split;
    Merge merge;
};

struct CurvatureParameters
{
    struct Split
    {
        float minimalCurvatureThreshold = 2.0F;
    };
    struct Merge
    {
        float maximalCurvatureThreshold = 0.6F;
    };
    Split split;
    Merge merge;
};

struct InterfaceParameters
{
    struct Split
    {
        float distanceRatioThreshold = 0.05F;
    };
    struct Merge
    {
        float distanceRatioThreshold = 0.15F;
    };
    Split split;
    Merge merge;
};

struct SplittingParameters
{
    float epsilon = 0.65F;
    float alpha = 0.70F;
    float centerMassRatio = 0.077F;
    float vertexMassRatio = 0.077F;
};

// Extension: moving spherical refinement zone (Criterion::Zone).
struct ZoneParameters
{
    vec3 center {};
    float radius = 0.0F;
    float pathRadius = 0.0F;
    float pathPeriod = 0.0F;
};

struct RefinementParameters
{
    enum class Criterion
    {
        Velocity,
        Interface,
        Vorticity,
        Zone  // extension
    };

    bool enabled = false;
    float minMassRatio = 0.3F;
    float maxMassRatio = 8.0F;
    uint32_t maxParticleCount = 1000000;
    float maxBatchRatio = 0.2F;
    uint32_t initialCooldown = 10;
    uint32_t cooldown = 10;
    Criterion criterionType = Criterion::Velocity;

    SplittingParameters splitting;
    VelocityParameters velocity;

    VorticityParameters vorticity;
    CurvatureParameters curvature;
    InterfaceParameters interfaceParameters;
    ZoneParameters zone;  // extension

    [[nodiscard]] asph_refinement_parameters toC() const
    {
        asph_refinement_parameters r {};
        detail::check(asph_default_refinement_parameters(&r), "default refinement parameters");
        r.enabled = enabled ? 1U : 0U;
        r.min_mass_ratio = minMassRatio;
        r.max_mass_ratio = maxMassRatio;
        r.max_particle_count = maxParticleCount;
        r.max_batch_ratio = maxBatchRatio;
        r.initial_cooldown = initialCooldown;
        r.cooldown = cooldown;
        r.criterion_type = static_cast<uint32_t>(criterionType);
        r.splitting = {splitting.epsilon, splitting.alpha, splitting.centerMassRatio, splitting.vertexMassRatio};
        r.velocity = {velocity.split.minimalSpeedThreshold, velocity.merge.maximalSpeedThreshold};
        r.vorticity = {vorticity.split.minimalVorticityThreshold, vorticity.merge.maximalVorticityThreshold};
        r.curvature = {curvature.split.minimalCurvatureThreshold, curvature.merge.maximalCurvatureThreshold};
        r.interface_ratio = {interfaceParameters.split.distanceRatioThreshold,
                             interfaceParameters.merge.distanceRatioThreshold};
        r.zone = {detail::toC(zone.center), zone.radius, zone.pathRadius, zone.pathPeriod};
        return r;
    }
};

}  // namespace refinement

namespace physics
{
class StaticBoundaryDomain;
}

class Simulation
{
public:
    // Simulation::Parameters (Simulation.cuh:64-109) + extensions.
    struct Parameters
    {
        struct Domain
        {
            vec3 min = {-1.F, -1.F, -1.F};
            vec3 max = {1.F, 1.F, 1.F};
            float friction = 0.01F;

            [[nodiscard]] vec3 getTranslation() const noexcept
            {
                return {(max.x + min.x) / 2.F, (max.y + min.y) / 2.F, (max.z + min.z) / 2.F};
            }
            [[nodiscard]] vec3 getScale() const noexcept { return {max.x - min.x, max.y - min.y, max.z - min.z}; }
            [[nodiscard]] float getVolume() const noexcept
            {
                const vec3 s = getScale();
                return s.x * s.y * s.z;
            }
            static Domain fromTransform(vec3 translation, vec3 scale)
            {
                Domain d;
                d.min = {translation.x - scale.x / 2.F, translation.y - scale.y / 2.F, translation.z - scale.z / 2.F};
                d.max = {translation.x + scale.x / 2.F, translation.y + scale.y / 2.F, translation.z + scale.z / 2.F};
                return d;
            }
            bool operator==(const Domain& o) const
            {
                return min.x == o.min.x && min.y == o.min.y && min.z == o.min.z && max.x == o.max.x &&
                       max.y == o.max.y && max.z == o.max.z && friction == o.friction;
            }
            [[nodiscard]] asph_domain toC() const { return {detail::toC(min), detail::toC(max), friction}; }
        };

        Domain domain;
        vec3 gravity {};
        float speedOfSound = 0.F;
        float restDensity = 0.F;
        float restitution = 0.F;
        float viscosityConstant = 0.F;
        float surfaceTensionConstant = 0.F;
        float maxVelocity = 0.F;
        float baseSmoothingRadius = 0.F;
        float baseParticleRadius = 0.F;
        float baseParticleMass = 0.F;

        uint32_t threadsPerBlock = 256;

        // extensions (defaults give the reference's behaviour)
        float taitGamma = 5.0F;
        float sloshAmplitude = 0.F;
        float sloshPeriod = 0.F;
        uint32_t flags = 0;  // ASPH_FLAG_*
        float laminarViscosity = 0.F;  // Morris laminar term's nu (0 = off)

        [[nodiscard]] asph_parameters toC() const
        {
            asph_parameters p {};
            detail::check(asph_default_parameters(&p), "default parameters");
            p.domain = domain.toC();
            p.gravity = detail::toC(gravity);
            p.speed_of_sound = speedOfSound;
            p.rest_density = restDensity;
            p.restitution = restitution;
            p.viscosity_constant = viscosityConstant;
            p.surface_tension_constant = surfaceTensionConstant;
            p.max_velocity = maxVelocity;
            p.base_smoothing_radius = baseSmoothingRadius;
            p.base_particle_radius = baseParticleRadius;
            p.base_particle_mass = baseParticleMass;
            p.threads_per_block = threadsPerBlock;
            p.tait_gamma = taitGamma;
            p.slosh_amplitude = sloshAmplitude;
            p.slosh_period = sloshPeriod;
            p.flags = flags;
            p.laminar_viscosity = laminarViscosity;
            return p;
        }
    };

    // Simulation::DensityInfo (Simulation.cuh:111-120).
    struct DensityInfo
    {
        float restDensity;
        float minDensity;
        float maxDensity;
        float averageDensity;
        uint32_t underDensityCount;
        uint32_t normalDensityCount;
        uint32_t overDensityCount;
    };

    explicit Simulation(asph_simulation* handle) : handle_(handle) {}
    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;
    virtual ~Simulation() { (void)asph_destroy(handle_); }

    // Simulation.cuh:124-137. update() is synchronous, as in the reference.
    virtual void update(float deltaTime) { detail::check(asph_update(handle_, deltaTime), "update"); }
    virtual void updateDomain(const Parameters::Domain& domain, const physics::StaticBoundaryDomain& boundaryDomain);

    [[nodiscard]] virtual uint32_t getFluidParticlesCount() const
    {
        uint32_t n = 0;
        detail::check(asph_fluid_count(handle_, &n), "getFluidParticlesCount");
        return n;
    }
    [[nodiscard]] virtual uint32_t getBoundaryParticlesCount() const
    {
        uint32_t n = 0;
        detail::check(asph_boundary_count(handle_, &n), "getBoundaryParticlesCount");
        return n;
    }
    [[nodiscard]] virtual float calculateAverageNeighborCount()
    {
        float v = 0.F;
        detail::check(asph_average_neighbor_count(handle_, &v), "calculateAverageNeighborCount");
        return v;
    }
    [[nodiscard]] virtual DensityInfo getDensityInfo(float threshold)
    {
        asph_density_info d {};
        detail::check(asph_get_density_info(handle_, threshold, &d), "getDensityInfo");
        return {d.rest_density,         d.min_density,          d.max_density,         d.average_density,
                d.under_density_count, d.normal_density_count, d.over_density_count};
    }
    virtual void enableAdaptiveRefinement()
    {
        detail::check(asph_enable_adaptive_refinement(handle_), "enableAdaptiveRefinement");
    }

    // Particle-buffer access (replaces the imported renderer buffers): device
    // pointers valid until the next call that changes the simulation.
    [[nodiscard]] FluidParticlesData fluidParticles()
    {
        asph_fluid_view v {};
        detail::check(asph_get_fluid_view(handle_, &v), "fluidParticles");
        return {reinterpret_cast<vec4*>(v.positions), reinterpret_cast<vec4*>(v.velocities),
                reinterpret_cast<vec4*>(v.accelerations), v.densities, v.radii, v.smoothing_radii, v.masses,
                v.particle_count};
    }
    // The boundary particles (positions, colours and radii are the imported
    // boundary buffers when the simulation was created with them).
    [[nodiscard]] BoundaryParticlesData boundaryParticles()
    {
        asph_boundary_view v {};
        detail::check(asph_get_boundary_view(handle_, &v), "boundaryParticles");
        return {reinterpret_cast<vec4*>(v.positions), reinterpret_cast<vec4*>(v.colors), v.psi_values,
                v.viscosity_coefficients, v.radii, v.particle_count};
    }
    // Host download of the positions (and optionally velocities and densities).
    void copyFluid(std::vector<vec4>* positions, std::vector<vec4>* velocities = nullptr,
                   std::vector<float>* densities = nullptr)
    {
        const uint32_t n = getFluidParticlesCount();
        uint32_t mask = 0;
        if (positions != nullptr)
        {
            positions->resize(n);
            mask |= ASPH_FIELD_POSITIONS;
        }
        if (velocities != nullptr)
        {
            velocities->resize(n);
            mask |= ASPH_FIELD_VELOCITIES;
        }
        if (densities != nullptr)
        {
            densities->resize(n);
            mask |= ASPH_FIELD_DENSITIES;
        }
        detail::check(asph_copy_fluid(handle_, mask,
                                      positions ? reinterpret_cast<asph_vec4*>(positions->data()) : nullptr,
                                      velocities ? reinterpret_cast<asph_vec4*>(velocities->data()) : nullptr,
                                      nullptr, densities ? densities->data() : nullptr, nullptr, nullptr, nullptr),
                      "copyFluid");
    }

    [[nodiscard]] asph_simulation* handle() const noexcept { return handle_; }

private:
    asph_simulation* handle_;
};

namespace physics
{

// StaticBoundaryDomain (StaticBoundaryDomain.cuh:14-50); generation runs on
// the host inside the library.
class StaticBoundaryDomain
{
public:
    struct Particle
    {
        vec4 position;
        float psi;
    };
    static_assert(sizeof(Particle) == sizeof(asph_boundary_particle), "Particle layout");

    static StaticBoundaryDomain generate(const Simulation::Parameters::Domain& bounds, float particleSpacing,
                                         float fluidRestDensity, float fluidSmoothingRadius)
    {
        const asph_domain d = bounds.toC();
        uint32_t n = 0;
        detail::check(asph_boundary_generate(&d, particleSpacing, fluidRestDensity, fluidSmoothingRadius, nullptr, &n),
                      "StaticBoundaryDomain::generate");
        std::vector<Particle> particles(n);
        detail::check(asph_boundary_generate(&d, particleSpacing, fluidRestDensity, fluidSmoothingRadius,
                                             reinterpret_cast<asph_boundary_particle*>(particles.data()), &n),
                      "StaticBoundaryDomain::generate");
        return StaticBoundaryDomain(std::move(particles));
    }

    [[nodiscard]] const std::vector<Particle>& getParticles() const { return particles_; }
    [[nodiscard]] uint32_t getParticleCount() const { return static_cast<uint32_t>(particles_.size()); }

private:
    explicit StaticBoundaryDomain(std::vector<Particle> p) : particles_(std::move(p)) {}
    std::vector<Particle> particles_;
};

}  // namespace physics

inline void Simulation::updateDomain(const Parameters::Domain& domain,
                                     const physics::StaticBoundaryDomain& boundaryDomain)
{
    const asph_domain d = domain.toC();
    const auto& b = boundaryDomain.getParticles();
    detail::check(asph_update_domain(handle_, &d, reinterpret_cast<const asph_boundary_particle*>(b.data()),
                                     static_cast<uint32_t>(b.size())),
                  "updateDomain");
}

// createSimulation (Simulation.cuh:138-145, SimulationFactory.cu:14-40): the
// simulation keeps its fluid positions, velocities, densities and radii in
// fluidParticleMemory and writes the walls' positions, radii and colours
// into boundaryParticleMemory (SphSimulation.cu:287-333); the adaptive
// variant is chosen when refinementParams->enabled and then uses
// maxParticleCount as its capacity (AdaptiveSphSimulation.cu:37), so the
// fluid buffers must hold that many particles. The memory must outlive the
// simulation.
inline std::unique_ptr<Simulation> createSimulation(
    const Simulation::Parameters& parameters, const std::vector<vec4>& positions,
    const FluidParticlesDataImportedBuffer& fluidParticleMemory,
    const BoundaryParticlesDataImportedBuffer& boundaryParticleMemory,
    const physics::StaticBoundaryDomain& boundaryDomain,
    const std::optional<refinement::RefinementParameters>& refinementParams, uint32_t maxFluidParticleCapacity,
    uint32_t maxBoundaryParticleCapacity)
{
    const asph_parameters p = parameters.toC();
    asph_refinement_parameters r {};
    const asph_refinement_parameters* rp = nullptr;
    if (refinementParams.has_value())
    {
        r = refinementParams->toC();
        rp = &r;
    }
    const asph_fluid_buffers fluid {fluidParticleMemory.positions.buffer(), fluidParticleMemory.velocities.buffer(),
                                    fluidParticleMemory.densities.buffer(), fluidParticleMemory.radii.buffer()};
    const asph_boundary_buffers walls {boundaryParticleMemory.positions.buffer(),
                                      boundaryParticleMemory.radii.buffer(), boundaryParticleMemory.colors.buffer()};
    const auto& b = boundaryDomain.getParticles();
    asph_simulation* h = nullptr;
    detail::check(asph_create_imported(&p, reinterpret_cast<const asph_vec4*>(positions.data()),
                                       static_cast<uint32_t>(positions.size()), &fluid, &walls,
                                       reinterpret_cast<const asph_boundary_particle*>(b.data()),
                                       static_cast<uint32_t>(b.size()), rp, maxFluidParticleCapacity,
                                       maxBoundaryParticleCapacity, &h),
                  "createSimulation");
    return std::make_unique<Simulation>(h);
}

// createSimulation without imported memory (the simulation allocates and
// owns its particle memory); otherwise as above.
inline std::unique_ptr<Simulation> createSimulation(
    const Simulation::Parameters& parameters, const std::vector<vec4>& positions,
    const physics::StaticBoundaryDomain& boundaryDomain,
    const std::optional<refinement::RefinementParameters>& refinementParams, uint32_t maxFluidParticleCapacity,
    uint32_t maxBoundaryParticleCapacity)
{
    const asph_parameters p = parameters.toC();
    asph_refinement_parameters r {};
    const asph_refinement_parameters* rp = nullptr;
    if (refinementParams.has_value())
    {
        r = refinementParams->toC();
        rp = &r;
    }
    const auto& b = boundaryDomain.getParticles();
    asph_simulation* h = nullptr;
    detail::check(asph_create(&p, reinterpret_cast<const asph_vec4*>(positions.data()),
                              static_cast<uint32_t>(positions.size()),
                              reinterpret_cast<const asph_boundary_particle*>(b.data()),
                              static_cast<uint32_t>(b.size()), rp, maxFluidParticleCapacity,
                              maxBoundaryParticleCapacity, &h),
                  "createSimulation");
    return std::make_unique<Simulation>(h);
}

}  // namespace sph::cuda
