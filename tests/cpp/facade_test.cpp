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
        std::cerr << "usage: facade_test host | run <params.bin> <positions.bin> <steps> <dt> <out.bin>\n";
        return 2;
    }
    catch (const std::exception& e)
    {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
