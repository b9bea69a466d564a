// Headless runner: reads a configuration file in the reference application's
// JSON schema (app/utils/Serializers.hpp:62-136, Configuration.hpp:12-34),
// builds the application's default scene (App.cpp:279-339: a particle grid
// filling the domain minus a 0.15 m margin, static walls from
// StaticBoundaryDomain::generate with spacing 2 r), and steps it through the
// C++ interface with the application's fixed time step (App.cpp:230), printing
// the diagnostics the application shows (average neighbour count, density
// info) as JSON lines.
//
//   asph_run CONFIG [--steps N] [--dt S] [--report-every N] [--dump FILE]
//   asph_run CONFIG --check        resolved configuration, no device work
//   asph_run --template FILE       configuration with every default
//
// Extension fields (not in the reference schema, optional): in
// simulationParameters "taitGamma", "sloshAmplitude", "sloshPeriod", "flags",
// "laminarViscosity";
// in refinementParameters a "zone" object and criterionType "zone".
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <optional>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/sph/Simulation.hpp"
#include "json.hpp"

using sph::cuda::Simulation;
using sph::cuda::vec3;
using sph::cuda::vec4;
namespace ref = sph::cuda::refinement;

namespace
{

struct Config
{
    uint32_t count[3] = {40, 40, 25};  // InitialParameters::particleCount default
    Simulation::Parameters sim;
    std::optional<ref::RefinementParameters> refinement;
};

vec3 readVec3(const asph::json::Value& v) { return {v.at("x").f32(), v.at("y").f32(), v.at("z").f32()}; }

Config readConfig(const asph::json::Value& root)
{
    Config c;
    // App::run requires both (App.cpp:127-128)
    const auto& ip = root.at("initialParameters");
    const auto& pc = ip.at("particleCount");
    c.count[0] = static_cast<uint32_t>(pc.at("x").u64());
    c.count[1] = static_cast<uint32_t>(pc.at("y").u64());
    c.count[2] = static_cast<uint32_t>(pc.at("z").u64());
    const auto& sp = root.at("simulationParameters");
    auto& p = c.sim;
    p.domain.min = readVec3(sp.at("domain").at("min"));
    p.domain.max = readVec3(sp.at("domain").at("max"));
    p.domain.friction = sp.at("domain").at("friction").f32();
    p.gravity = readVec3(sp.at("gravity"));
    p.speedOfSound = sp.at("speedOfSound").f32();
    p.restDensity = sp.at("restDensity").f32();
    p.restitution = sp.at("restitution").f32();
    p.viscosityConstant = sp.at("viscosityConstant").f32();
    p.surfaceTensionConstant = sp.at("surfaceTensionConstant").f32();
    p.maxVelocity = sp.at("maxVelocity").f32();
    p.baseSmoothingRadius = sp.at("baseSmoothingRadius").f32();
    p.baseParticleRadius = sp.at("baseParticleRadius").f32();
    p.baseParticleMass = sp.at("baseParticleMass").f32();
    p.threadsPerBlock = static_cast<uint32_t>(sp.at("threadsPerBlock").u64());
    if (sp.has("taitGamma"))
    {
        p.taitGamma = sp.at("taitGamma").f32();
    }
    if (sp.has("sloshAmplitude"))
    {
        p.sloshAmplitude = sp.at("sloshAmplitude").f32();
    }
    if (sp.has("sloshPeriod"))
    {
        p.sloshPeriod = sp.at("sloshPeriod").f32();
    }
    if (sp.has("flags"))
    {
        p.flags = static_cast<uint32_t>(sp.at("flags").u64());
    }
    if (sp.has("laminarViscosity"))
    {
        p.laminarViscosity = sp.at("laminarViscosity").f32();
    }
    if (root.has("refinementParameters") && !root.at("refinementParameters").isNull())
    {
        const auto& rp = root.at("refinementParameters");
        ref::RefinementParameters r;
        r.enabled = rp.at("enabled").flag();
        r.minMassRatio = rp.at("minMassRatio").f32();
        r.maxMassRatio = rp.at("maxMassRatio").f32();
        r.maxParticleCount = static_cast<uint32_t>(rp.at("maxParticleCount").u64());
        r.maxBatchRatio = rp.at("maxBatchRatio").f32();
        r.initialCooldown = static_cast<uint32_t>(rp.at("initialCooldown").u64());
        r.cooldown = static_cast<uint32_t>(rp.at("cooldown").u64());
        const std::string crit = rp.at("criterionType").str();
        if (crit == "velocity")
        {
            r.criterionType = ref::RefinementParameters::Criterion::Velocity;
        }
        else if (crit == "interface")
        {
            r.criterionType = ref::RefinementParameters::Criterion::Interface;
        }
        else if (crit == "vorticity")
        {
            r.criterionType = ref::RefinementParameters::Criterion::Vorticity;
        }
        else if (crit == "zone")
        {
            r.criterionType = ref::RefinementParameters::Criterion::Zone;
        }
        else
        {
            throw std::runtime_error("unknown criterionType \"" + crit + "\"");
        }
        const auto& sl = rp.at("splitting");
        r.splitting.epsilon = sl.at("epsilon").f32();
        r.splitting.alpha = sl.at("alpha").f32();
        r.splitting.centerMassRatio = sl.at("centerMassRatio").f32();
        r.splitting.vertexMassRatio = sl.at("vertexMassRatio").f32();
        r.velocity.split.minimalSpeedThreshold = rp.at("velocity").at("split").at("minimalSpeedThreshold").f32();
        r.velocity.merge.maximalSpeedThreshold = rp.at("velocity").at("merge").at("maximalSpeedThreshold").f32();
        r.vorticity.split.minimalVorticityThreshold =
            rp.at("vorticity").at("split").at("minimalVorticityThreshold").f32();
        r.vorticity.merge.maximalVorticityThreshold =
            rp.at("vorticity").at("merge").at("maximalVorticityThreshold").f32();
        r.curvature.split.minimalCurvatureThreshold =
            rp.at("curvature").at("split").at("minimalCurvatureThreshold").f32();
        r.curvature.merge.maximalCurvatureThreshold =
            rp.at("curvature").at("merge").at("maximalCurvatureThreshold").f32();
        r.interfaceParameters.split.distanceRatioThreshold =
            rp.at("interfaceParameters").at("split").at("distanceRatioThreshold").f32();
        r.interfaceParameters.merge.distanceRatioThreshold =
            rp.at("interfaceParameters").at("merge").at("distanceRatioThreshold").f32();
        if (rp.has("zone"))
        {
            const auto& z = rp.at("zone");
            r.zone.center = readVec3(z.at("center"));
            r.zone.radius = z.at("radius").f32();
            r.zone.pathRadius = z.has("pathRadius") ? z.at("pathRadius").f32() : 0.F;
            r.zone.pathPeriod = z.has("pathPeriod") ? z.at("pathPeriod").f32() : 0.F;
        }
        c.refinement = r;
    }
    return c;
}

std::string num(float v)
{
    char b[64];
    std::snprintf(b, sizeof(b), "%.9g", static_cast<double>(v));
    return b;
}

std::string vec(const vec3& v) { return "{\"x\": " + num(v.x) + ", \"y\": " + num(v.y) + ", \"z\": " + num(v.z) + "}"; }

std::string writeConfig(const Config& c)
{
    const auto& p = c.sim;
    std::ostringstream o;
    o << "{\n    \"initialParameters\": {\"particleCount\": {\"x\": " << c.count[0] << ", \"y\": " << c.count[1]
      << ", \"z\": " << c.count[2] << "}},\n";
    o << "    \"simulationParameters\": {\n"
      << "        \"domain\": {\"min\": " << vec(p.domain.min) << ", \"max\": " << vec(p.domain.max)
      << ", \"friction\": " << num(p.domain.friction) << "},\n"
      << "        \"gravity\": " << vec(p.gravity) << ",\n"
      << "        \"speedOfSound\": " << num(p.speedOfSound) << ",\n"
      << "        \"restDensity\": " << num(p.restDensity) << ",\n"
      << "        \"restitution\": " << num(p.restitution) << ",\n"
      << "        \"viscosityConstant\": " << num(p.viscosityConstant) << ",\n"
      << "        \"surfaceTensionConstant\": " << num(p.surfaceTensionConstant) << ",\n"
      << "        \"maxVelocity\": " << num(p.maxVelocity) << ",\n"
      << "        \"baseSmoothingRadius\": " << num(p.baseSmoothingRadius) << ",\n"
      << "        \"baseParticleRadius\": " << num(p.baseParticleRadius) << ",\n"
      << "        \"baseParticleMass\": " << num(p.baseParticleMass) << ",\n"
      << "        \"threadsPerBlock\": " << p.threadsPerBlock << ",\n"
      << "        \"taitGamma\": " << num(p.taitGamma) << ",\n"
      << "        \"sloshAmplitude\": " << num(p.sloshAmplitude) << ",\n"
      << "        \"sloshPeriod\": " << num(p.sloshPeriod) << ",\n"
      << "        \"flags\": " << p.flags << ",\n"
      << "        \"laminarViscosity\": " << num(p.laminarViscosity) << "\n    },\n";
    if (!c.refinement)
    {
        o << "    \"refinementParameters\": null\n}\n";
        return o.str();
    }
    const auto& r = *c.refinement;
    const char* crit[] = {"velocity", "interface", "vorticity", "zone"};
    o << "    \"refinementParameters\": {\n"
      << "        \"enabled\": " << (r.enabled ? "true" : "false") << ",\n"
      << "        \"minMassRatio\": " << num(r.minMassRatio) << ",\n"
      << "        \"maxMassRatio\": " << num(r.maxMassRatio) << ",\n"
      << "        \"maxParticleCount\": " << r.maxParticleCount << ",\n"
      << "        \"maxBatchRatio\": " << num(r.maxBatchRatio) << ",\n"
      << "        \"initialCooldown\": " << r.initialCooldown << ",\n"
      << "        \"cooldown\": " << r.cooldown << ",\n"
      << "        \"criterionType\": \"" << crit[static_cast<int>(r.criterionType)] << "\",\n"
      << "        \"splitting\": {\"epsilon\": " << num(r.splitting.epsilon) << ", \"alpha\": " << num(r.splitting.alpha)
      << ", \"centerMassRatio\": " << num(r.splitting.centerMassRatio)
      << ", \"vertexMassRatio\": " << num(r.splitting.vertexMassRatio) << "},\n"
      << "        \"velocity\": {\"split\": {\"minimalSpeedThreshold\": " << num(r.velocity.split.minimalSpeedThreshold)
      << "}, \"merge\": {\"maximalSpeedThreshold\": " << num(r.velocity.merge.maximalSpeedThreshold) << "}},\n"
      << "        \"vorticity\": {\"split\": {\"minimalVorticityThreshold\": "
      << num(r.vorticity.split.minimalVorticityThreshold) << "}, \"merge\": {\"maximalVorticityThreshold\": "
      << num(r.vorticity.merge.maximalVorticityThreshold) << "}},\n"
      << "        \"curvature\": {\"split\": {\"minimalCurvatureThreshold\": "
      << num(r.curvature.split.minimalCurvatureThreshold) << "}, \"merge\": {\"maximalCurvatureThreshold\": "
      << num(r.curvature.merge.maximalCurvatureThreshold) << "}},\n"
      << "        \"interfaceParameters\": {\"split\": {\"distanceRatioThreshold\": "
      << num(r.interfaceParameters.split.distanceRatioThreshold) << "}, \"merge\": {\"distanceRatioThreshold\": "
      << num(r.interfaceParameters.merge.distanceRatioThreshold) << "}},\n"
      << "        \"zone\": {\"center\": " << vec(r.zone.center) << ", \"radius\": " << num(r.zone.radius)
      << ", \"pathRadius\": " << num(r.zone.pathRadius) << ", \"pathPeriod\": " << num(r.zone.pathPeriod)
      << "}\n    }\n}\n";
    return o.str();
}

// App::createParticlesInGrid (App.cpp:312-339), in FP32 as the application
// computes it: x outer, z innermost.
std::vector<vec4> particleGrid(const Config& c)
{
    const auto& d = c.sim.domain;
    constexpr float margin = 0.15F;
    const float size[3] = {d.max.x - d.min.x, d.max.y - d.min.y, d.max.z - d.min.z};
    float spacing[3];
    for (int a = 0; a < 3; ++a)
    {
        const float avail = size[a] - 2.0F * margin;
        spacing[a] = avail / static_cast<float>(c.count[a] - 1U);
    }
    std::vector<vec4> out;
    out.reserve(static_cast<size_t>(c.count[0]) * c.count[1] * c.count[2]);
    const float bx = d.min.x + margin;
    const float by = d.min.y + margin;
    const float bz = d.min.z + margin;
    for (uint32_t i = 0; i < c.count[0]; ++i)
    {
        for (uint32_t j = 0; j < c.count[1]; ++j)
        {
            for (uint32_t k = 0; k < c.count[2]; ++k)
            {
                vec4 p;
                p.x = bx + spacing[0] * static_cast<float>(i);
                p.y = by + spacing[1] * static_cast<float>(j);
                p.z = bz + spacing[2] * static_cast<float>(k);
                p.w = 0.F;
                out.push_back(p);
            }
        }
    }
    return out;
}

std::string readFile(const std::string& path)
{
    std::ifstream f(path);
    if (!f)
    {
        throw std::runtime_error("cannot open " + path);
    }
    std::ostringstream s;
    s << f.rdbuf();
    return s.str();
}

void report(Simulation& sim, uint64_t step, double t)
{
    const auto info = sim.getDensityInfo(0.1F);
    std::printf(
        "{\"step\": %llu, \"time\": %.6g, \"particles\": %u, \"average_neighbors\": %.6g, \"density\": {\"rest\": %.6g, "
        "\"min\": %.6g, \"max\": %.6g, \"average\": %.6g, \"under\": %u, \"normal\": %u, \"over\": %u}}\n",
        static_cast<unsigned long long>(step), t, sim.getFluidParticlesCount(),
        static_cast<double>(sim.calculateAverageNeighborCount()), static_cast<double>(info.restDensity),
        static_cast<double>(info.minDensity), static_cast<double>(info.maxDensity),
        static_cast<double>(info.averageDensity), info.underDensityCount, info.normalDensityCount,
        info.overDensityCount);
    std::fflush(stdout);
}

int usage()
{
    std::cerr << "usage: asph_run CONFIG [--steps N] [--dt S] [--report-every N] [--dump FILE] [--check]\n"
                 "       asph_run --template FILE\n";
    return 2;
}

}  // namespace

int main(int argc, char** argv)
{
    try
    {
        if (argc == 3 && std::strcmp(argv[1], "--template") == 0)
        {
            Config c;  // the reference's struct defaults (Simulation.cuh, RefinementParameters.cuh)
            c.sim.domain = Simulation::Parameters::Domain {};
            c.refinement = ref::RefinementParameters {};
            std::ofstream(argv[2]) << writeConfig(c);
            return 0;
        }
        if (argc < 2)
        {
            return usage();
        }
        uint64_t steps = 1000;
        float dt = 0.0002F;  // App.cpp:230
        uint64_t every = 100;
        std::string dump;
        bool check = false;
        for (int a = 2; a < argc; ++a)
        {
            const std::string k = argv[a];
            if (k == "--check")
            {
                check = true;
            }
            else if (a + 1 < argc && k == "--steps")
            {
                steps = std::stoull(argv[++a]);
            }
            else if (a + 1 < argc && k == "--dt")
            {
                dt = std::stof(argv[++a]);
            }
            else if (a + 1 < argc && k == "--report-every")
            {
                every = std::stoull(argv[++a]);
            }
            else if (a + 1 < argc && k == "--dump")
            {
                dump = argv[++a];
            }
            else
            {
                return usage();
            }
        }
        const Config cfg = readConfig(asph::json::parse(readFile(argv[1])));
        if (check)
        {
            std::cout << writeConfig(cfg);
            return 0;
        }
        const auto particles = particleGrid(cfg);
        const auto& p = cfg.sim;
        const auto boundary = sph::cuda::physics::StaticBoundaryDomain::generate(
            p.domain, p.baseParticleRadius * 2.F, p.restDensity, p.baseSmoothingRadius);
        // capacities as App.cpp:151-172
        const bool adaptive = cfg.refinement.has_value() && cfg.refinement->enabled;
        const uint32_t fluidCap = adaptive ? cfg.refinement->maxParticleCount : static_cast<uint32_t>(particles.size());
        auto sim = sph::cuda::createSimulation(p, particles, boundary, cfg.refinement, fluidCap,
                                               boundary.getParticleCount() * 10U);
        report(*sim, 0, 0.0);
        const auto t0 = std::chrono::steady_clock::now();
        uint64_t updates = 0;
        for (uint64_t s = 1; s <= steps; ++s)
        {
            updates += sim->getFluidParticlesCount();
            sim->update(dt);
            if (every != 0 && (s % every == 0 || s == steps))
            {
                report(*sim, s, static_cast<double>(s) * static_cast<double>(dt));
            }
        }
        const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("{\"summary\": {\"steps\": %llu, \"particles\": %u, \"boundary\": %u, \"wall_s\": %.6g, "
                    "\"particle_updates_per_s\": %.6g}}\n",
                    static_cast<unsigned long long>(steps), sim->getFluidParticlesCount(),
                    sim->getBoundaryParticlesCount(), wall, wall > 0 ? static_cast<double>(updates) / wall : 0.0);
        if (!dump.empty())
        {
            std::vector<vec4> x;
            std::vector<vec4> v;
            std::vector<float> rho;
            sim->copyFluid(&x, &v, &rho);
            const uint32_t n = static_cast<uint32_t>(x.size());
            std::ofstream f(dump, std::ios::binary);
            f.write(reinterpret_cast<const char*>(&n), sizeof(n));
            f.write(reinterpret_cast<const char*>(x.data()), static_cast<std::streamsize>(n * sizeof(vec4)));
            f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(n * sizeof(vec4)));
            f.write(reinterpret_cast<const char*>(rho.data()), static_cast<std::streamsize>(n * sizeof(float)));
        }
        return 0;
    }
    catch (const std::exception& e)
    {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
}
