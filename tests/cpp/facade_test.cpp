// Test program of the C++ interface (include/sph/Simulation.hpp), written the
// way the reference's application drives the reference (App.cpp:143-172,
// 212-233): generate the static boundary, create the simulation, update,
// query.
//
//   facade_test host
//       checks that need no GPU: defaults of the parameter structs against
//       the library's, Domain helpers, boundary generation, and that invalid
//       parameters surface as std::runtime_error.
//   facade_test run <params.bin> <positions.bin> <steps> <dt> <out.bin>
//       reads an asph_parameters record and N float4 positions, creates the
//       simulation through the C++ interface, runs <steps> updates and writes
//       count, positions, velocities and densities to <out.bin>.
//   facade_test imported <params.bin> <positions.bin> <steps> <dt> <out.bin> [<refinement.bin>]
//       the same through the reference's 8-argument createSimulation with
//       imported particle memory, the call written as App.cpp:154-172 writes
//       it; the results are read back from the imported buffers (what the
//       renderer draws): count, positions, velocities, densities, then the
//       boundary count, positions, radii and colours.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sph/Simulation.hpp"

using namespace sph::cuda;

namespace
{

int failures = 0;

void expect(bool ok, const std::string& what)
{
    if (!ok)
    {
        std::cerr << "FAIL: " << what << "\n";
        ++failures;
    }
}

std::vector<char> readFile(const char* path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
    {
        throw std::runtime_error(std::string("cannot open ") + path);
    }
    return {std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>()};
}

Simulation::Parameters fromC(const asph_parameters& c)
{
    Simulation::Parameters p;
    p.domain.min = {c.domain.min.x, c.domain.min.y, c.domain.min.z};
    p.domain.max = {c.domain.max.x, c.domain.max.y, c.domain.max.z};
    p.domain.friction = c.domain.friction;
    p.gravity = {c.gravity.x, c.gravity.y, c.gravity.z};
    p.speedOfSound = c.speed_of_sound;
    p.restDensity = c.rest_density;
    p.restitution = c.restitution;
    p.viscosityConstant = c.viscosity_constant;
    p.surfaceTensionConstant = c.surface_tension_constant;
    p.maxVelocity = c.max_velocity;
    p.baseSmoothingRadius = c.base_smoothing_radius;
    p.baseParticleRadius = c.base_particle_radius;
    p.baseParticleMass = c.base_particle_mass;
    p.threadsPerBlock = c.threads_per_block;
    p.taitGamma = c.tait_gamma;
    p.sloshAmplitude = c.slosh_amplitude;
    p.sloshPeriod = c.slosh_period;
    p.flags = c.flags;
    return p;
}

int host()
{
    // RefinementParameters defaults (RefinementParameters.cuh) == the library's
    const refinement::RefinementParameters r;
    asph_refinement_parameters lib {};
    expect(asph_default_refinement_parameters(&lib) == ASPH_OK, "default refinement parameters");
    const asph_refinement_parameters mine = r.toC();
    expect(std::memcmp(&mine, &lib, sizeof(lib)) == 0, "RefinementParameters defaults match the library");
    expect(!r.enabled && r.maxParticleCount == 1000000 && r.cooldown == 10 && r.initialCooldown == 10,
           "RefinementParameters scalar defaults");

    Simulation::Parameters p;
    asph_parameters c {};
    expect(asph_default_parameters(&c) == ASPH_OK, "default parameters");
    expect(p.threadsPerBlock == c.threads_per_block && p.taitGamma == c.tait_gamma && p.flags == c.flags,
           "Parameters extension defaults match the library");
    expect(p.domain.min.x == -1.F && p.domain.max.z == 1.F && p.domain.friction == 0.01F, "Domain defaults");

    const auto d = Simulation::Parameters::Domain::fromTransform({0.5F, 0.3F, 0.25F}, {1.0F, 0.6F, 0.5F});
    expect(d.min.x == 0.F && d.min.y == 0.F && d.min.z == 0.F, "fromTransform min");
    expect(std::fabs(d.getVolume() - 0.3F) < 1e-6F, "getVolume");
    expect(d.getTranslation().y == 0.3F && d.getScale().z == 0.5F, "translation and scale");

    // the 1.0 x 0.6 x 0.5 m box at spacing 0.01 m (r = 0.005 m) and h = 0.013 m
    const auto b = physics::StaticBoundaryDomain::generate(d, 0.01F, 1000.F, 0.013F);
    expect(b.getParticleCount() == 37482, "StaticBoundaryDomain::generate count");
    expect(b.getParticles().front().psi > 0.F, "psi positive");

    // invalid parameters are rejected before any device work
    Simulation::Parameters bad = p;
    bad.baseSmoothingRadius = -1.F;
    bool threw = false;
    try
    {
        const std::vector<vec4> one {{0.F, 0.F, 0.F, 0.F}};
        (void)createSimulation(bad, one, b, std::nullopt, 1, b.getParticleCount());
    }
    catch (const std::runtime_error& e)
    {
        threw = std::string(e.what()).find("smoothing") != std::string::npos;
    }
    expect(threw, "invalid parameters throw std::runtime_error naming the problem");

    if (failures == 0)
    {
        std::cout << "host ok\n";
    }
    return failures == 0 ? 0 : 1;
}

// The renderer's particle systems (App.cpp's _api->initializeParticleSystem /
// initializeBoundaryParticleSystem allocate Vulkan buffers and import them
// into the simulation): device allocations standing in for them.
struct ParticleSystem
{
    explicit ParticleSystem(uint32_t capacity)
        : positions(allocateBuffer(capacity * sizeof(vec4))), velocities(allocateBuffer(capacity * sizeof(vec4))),
          densities(allocateBuffer(capacity * sizeof(float))), radii(allocateBuffer(capacity * sizeof(float)))
    {
    }
    std::unique_ptr<ImportedMemory> positions, velocities, densities, radii;
};

struct BoundaryParticleSystem
{
    explicit BoundaryParticleSystem(uint32_t capacity)
        : positions(allocateBuffer(capacity * sizeof(vec4))), radii(allocateBuffer(capacity * sizeof(float))),
          colors(allocateBuffer(capacity * sizeof(vec4)))
    {
    }
    std::unique_ptr<ImportedMemory> positions, radii, colors;
};

class Api
{
public:
    FluidParticlesDataImportedBuffer initializeParticleSystem(uint32_t capacity)
    {
        fluid = std::make_unique<ParticleSystem>(capacity);
        return {*fluid->positions, *fluid->velocities, *fluid->densities, *fluid->radii};
    }
    BoundaryParticlesDataImportedBuffer initializeBoundaryParticleSystem(uint32_t capacity, bool /*render*/)
    {
        boundary = std::make_unique<BoundaryParticleSystem>(capacity);
        return {*boundary->positions, *boundary->radii, *boundary->colors};
    }
    std::unique_ptr<ParticleSystem> fluid;
    std::unique_ptr<BoundaryParticleSystem> boundary;
};

// the application's initial particle grid (ParticlesParameters::getScalarCount)
struct InitialParameters
{
    uint32_t count;
    [[nodiscard]] uint32_t getScalarCount() const { return count; }
};

struct Configuration
{
    std::optional<refinement::RefinementParameters> refinementParameters;
};

refinement::RefinementParameters refinementFromC(const asph_refinement_parameters& c)
{
    refinement::RefinementParameters r;
    r.enabled = c.enabled != 0U;
    r.minMassRatio = c.min_mass_ratio;
    r.maxMassRatio = c.max_mass_ratio;
    r.maxParticleCount = c.max_particle_count;
    r.maxBatchRatio = c.max_batch_ratio;
    r.initialCooldown = c.initial_cooldown;
    r.cooldown = c.cooldown;
    r.criterionType = static_cast<refinement::RefinementParameters::Criterion>(c.criterion_type);
    r.splitting.epsilon = c.splitting.epsilon;
    r.splitting.alpha = c.splitting.alpha;
    r.splitting.centerMassRatio = c.splitting.center_mass_ratio;
    r.splitting.vertexMassRatio = c.splitting.vertex_mass_ratio;
    r.velocity.split.minimalSpeedThreshold = c.velocity.split;
    r.velocity.merge.maximalSpeedThreshold = c.velocity.merge;
    r.vorticity.split.minimalVorticityThreshold = c.vorticity.split;
    r.vorticity.merge.maximalVorticityThreshold = c.vorticity.merge;
    r.curvature.split.minimalCurvatureThreshold = c.curvature.split;
    r.curvature.merge.maximalCurvatureThreshold = c.curvature.merge;
    r.interfaceParameters.split.distanceRatioThreshold = c.interface_ratio.split;
    r.interfaceParameters.merge.distanceRatioThreshold = c.interface_ratio.merge;
    r.zone.center = {c.zone.center.x, c.zone.center.y, c.zone.center.z};
    r.zone.radius = c.zone.radius;
    r.zone.pathRadius = c.zone.path_radius;
    r.zone.pathPeriod = c.zone.path_period;
    return r;
}

template <class T>
std::vector<T> download(const ImportedMemory& m, uint32_t n)
{
    std::vector<T> v(n);
    m.copyToHost(v.data(), n * sizeof(T));
    return v;
}

int imported(const char* paramsPath, const char* positionsPath, int steps, float dt, const char* outPath,
             const char* refinementPath)
{
    const auto pb = readFile(paramsPath);
    if (pb.size() != sizeof(asph_parameters))
    {
        throw std::runtime_error("parameters file has the wrong size");
    }
    asph_parameters c {};
    std::memcpy(&c, pb.data(), sizeof(c));
    const auto xb = readFile(positionsPath);
    std::vector<vec4> _particles(xb.size() / sizeof(vec4));
    std::memcpy(_particles.data(), xb.data(), _particles.size() * sizeof(vec4));
    Configuration config;
    if (refinementPath != nullptr)
    {
        const auto rb = readFile(refinementPath);
        asph_refinement_parameters rc {};
        if (rb.size() != sizeof(rc))
        {
            throw std::runtime_error("refinement file has the wrong size");
        }
        std::memcpy(&rc, rb.data(), sizeof(rc));
        config.refinementParameters = refinementFromC(rc);
    }
    const Simulation::Parameters _simulationParameters = fromC(c);
    const InitialParameters initialParameters {static_cast<uint32_t>(_particles.size())};
    Api api;
    Api* _api = &api;

    // App.cpp:143-172
    const auto particleSpacing = _simulationParameters.baseParticleRadius * 2;
    const auto boundaryDomain =
        physics::StaticBoundaryDomain::generate(_simulationParameters.domain,
                                                particleSpacing,
                                                _simulationParameters.restDensity,
                                                _simulationParameters.baseSmoothingRadius);

    const auto maxBoundaryCapacity = boundaryDomain.getParticleCount() * 10;

    // (App.cpp computes the fluid capacity with std::optional::and_then, C++23)
    const auto refinementCapacity = [](const std::optional<refinement::RefinementParameters>& r) {
        return r.has_value() && r->enabled ? std::optional<uint32_t>(r->maxParticleCount) : std::nullopt;
    };
    auto _simulation = createSimulation(
        _simulationParameters,
        _particles,
        _api->initializeParticleSystem(
            refinementCapacity(config.refinementParameters).value_or(initialParameters.getScalarCount())),
        _api->initializeBoundaryParticleSystem(maxBoundaryCapacity, true),
        boundaryDomain,
        config.refinementParameters,
        initialParameters.getScalarCount(),
        maxBoundaryCapacity);

    for (int k = 0; k < steps; ++k)
    {
        _simulation->update(dt);
    }
    const uint32_t n = _simulation->getFluidParticlesCount();
    const uint32_t nb = _simulation->getBoundaryParticlesCount();
    // the simulation's arrays are the imported buffers
    const FluidParticlesData view = _simulation->fluidParticles();
    const BoundaryParticlesData walls = _simulation->boundaryParticles();
    if (static_cast<void*>(view.positions) != api.fluid->positions->devicePointer() ||
        static_cast<void*>(view.velocities) != api.fluid->velocities->devicePointer() ||
        static_cast<void*>(view.densities) != api.fluid->densities->devicePointer() ||
        static_cast<void*>(view.radii) != api.fluid->radii->devicePointer() ||
        static_cast<void*>(walls.positions) != api.boundary->positions->devicePointer() ||
        static_cast<void*>(walls.radii) != api.boundary->radii->devicePointer() ||
        static_cast<void*>(walls.colors) != api.boundary->colors->devicePointer() || walls.particleCount != nb ||
        view.particleCount != n)
    {
        throw std::runtime_error("the simulation does not use the imported buffers");
    }
    const auto x = download<vec4>(*api.fluid->positions, n);
    const auto v = download<vec4>(*api.fluid->velocities, n);
    const auto rho = download<float>(*api.fluid->densities, n);
    const auto bx = download<vec4>(*api.boundary->positions, nb);
    const auto br = download<float>(*api.boundary->radii, nb);
    const auto bc = download<vec4>(*api.boundary->colors, nb);
    std::ofstream f(outPath, std::ios::binary);
    f.write(reinterpret_cast<const char*>(&n), sizeof(n));
    f.write(reinterpret_cast<const char*>(x.data()), static_cast<std::streamsize>(n * sizeof(vec4)));
    f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(n * sizeof(vec4)));
    f.write(reinterpret_cast<const char*>(rho.data()), static_cast<std::streamsize>(n * sizeof(float)));
    f.write(reinterpret_cast<const char*>(&nb), sizeof(nb));
    f.write(reinterpret_cast<const char*>(bx.data()), static_cast<std::streamsize>(nb * sizeof(vec4)));
    f.write(reinterpret_cast<const char*>(br.data()), static_cast<std::streamsize>(nb * sizeof(float)));
    f.write(reinterpret_cast<const char*>(bc.data()), static_cast<std::streamsize>(nb * sizeof(vec4)));
    std::cout << "imported ok: " << n << " particles, " << nb << " boundary particles, " << steps << " steps\n";
    return 0;
}

int run(const char* paramsPath, const char* positionsPath, int steps, float dt, const char* outPath)
{
    const auto pb = readFile(paramsPath);
    if (pb.size() != sizeof(asph_parameters))
    {
        throw std::runtime_error("parameters file has the wrong size");
    }
    asph_parameters c {};
    std::memcpy(&c, pb.data(), sizeof(c));
    const auto xb = readFile(positionsPath);
    std::vector<vec4> positions(xb.size() / sizeof(vec4));
    std::memcpy(positions.data(), xb.data(), positions.size() * sizeof(vec4));

    const Simulation::Parameters params = fromC(c);
    // App.cpp:142-151: wall spacing 2 r, boundary capacity 10x
    const float spacing = params.baseParticleRadius * 2.F;
    const auto boundary =
        physics::StaticBoundaryDomain::generate(params.domain, spacing, params.restDensity, params.baseSmoothingRadius);
    auto sim = createSimulation(params, positions, boundary, std::nullopt, static_cast<uint32_t>(positions.size()),
                                boundary.getParticleCount() * 10);
    for (int k = 0; k < steps; ++k)
    {
        sim->update(dt);
    }
    const uint32_t n = sim->getFluidParticlesCount();
    const FluidParticlesData view = sim->fluidParticles();
    if (view.particleCount != n || view.positions == nullptr)
    {
        throw std::runtime_error("fluidParticles view inconsistent");
    }
    std::vector<vec4> x;
    std::vector<vec4> v;
    std::vector<float> rho;
    sim->copyFluid(&x, &v, &rho);
    std::ofstream f(outPath, std::ios::binary);
    f.write(reinterpret_cast<const char*>(&n), sizeof(n));
    f.write(reinterpret_cast<const char*>(x.data()), static_cast<std::streamsize>(n * sizeof(vec4)));
    f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(n * sizeof(vec4)));
    f.write(reinterpret_cast<const char*>(rho.data()), static_cast<std::streamsize>(n * sizeof(float)));
    const auto info = sim->getDensityInfo(0.1F);
    std::cout << "run ok: " << n << " particles, " << steps << " steps, boundary "
              << sim->getBoundaryParticlesCount() << ", neighbours " << sim->calculateAverageNeighborCount()
              << ", density " << info.minDensity << ".." << info.maxDensity << "\n";
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    try
    {
        if (argc >= 2 && std::string(argv[1]) == "host")
        {
            return host();
        }
        if (argc == 7 && std::string(argv[1]) == "run")
        {
            return run(argv[2], argv[3], std::stoi(argv[4]), std::stof(argv[5]), argv[6]);
        }
        if ((argc == 7 || argc == 8) && std::string(argv[1]) == "imported")
        {
            return imported(argv[2], argv[3], std::stoi(argv[4]), std::stof(argv[5]), argv[6],
                            argc == 8 ? argv[7] : nullptr);
        }
        std::cerr << "usage: facade_test host | run|imported <params.bin> <positions.bin> <steps> <dt> <out.bin> "
                     "[<refinement.bin>]\n";
        return 2;
    }
    catch (const std::exception& e)
    {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
