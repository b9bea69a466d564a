/*
 * asph.h - C ABI of the MI355X-native per-step SPH particle pipeline.
 *
 * Every entry point returns an int status (ASPH_OK = 0). No C++ exception
 * crosses this boundary; asph_last_error() returns the message of the last
 * failure on the calling thread. All particle arrays are plain pointers and
 * element counts; vectors are 16-byte {x, y, z, w} records (w = 0), the
 * layout of glm::vec4 in the reference interface.
 *
 * The structs below mirror the reference's parameter PODs field for field
 * (reference: cuda/include/cuda/simulation/Simulation.cuh:64-120 and
 * cuda/include/cuda/refinement/RefinementParameters.cuh:8-104); fields marked
 * "extension" are appended after the mirrored prefix and have no counterpart
 * in the reference.
 */
#ifndef ASPH_H
#define ASPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASPH_OK 0
#define ASPH_ERR_INVALID 1   /* invalid argument or parameter */
#define ASPH_ERR_CAPACITY 2  /* a capacity or budget would be exceeded */
#define ASPH_ERR_DEVICE 3    /* a HIP runtime call failed */
#define ASPH_ERR_STATE 4     /* call not valid in the current state */

/* Largest fluid particle capacity of a simulation or batch pool (2^28 - 64):
   the neighbour loops address particle records by 32-bit byte offsets. */
#define ASPH_MAX_PARTICLE_CAPACITY 268435392u

typedef struct asph_vec3 { float x, y, z; } asph_vec3;
typedef struct asph_vec4 { float x, y, z, w; } asph_vec4;

/* Simulation::Parameters::Domain (Simulation.cuh:66-96). */
typedef struct asph_domain {
    asph_vec3 min;
    asph_vec3 max;
    float friction;
} asph_domain;

/* asph_parameters.flags */
#define ASPH_FLAG_FIX_CELL_START 1u /* build cell ranges without the reference's
                                      slot-0 omission (NeighborKernels.cu:27-36);
                                      default off = reference behaviour */
#define ASPH_FLAG_KERNEL_MERGED_H 2u /* merged smoothing length with the Wendland
                                       normalisation 21/(16 pi), so that a
                                       coincident equal pair keeps its h,
                                       instead of the reference's 16 pi/21
                                       (AdaptiveAlgorithm.cu:62-66), which
                                       grows h by (16 pi/21)^(2/3) = 1.79 per
                                       merge; default off = reference */

#define ASPH_FLAG_TOLERANCE_MATH 4u /* density and force sums over the neighbour
                                      lists with the hardware reciprocal square
                                      root and reciprocal, pressure and viscosity
                                      sharing one kernel-gradient factor: within
                                      1e-5 of the reference's FP32 results
                                      (forces relative to the sum of the pair
                                      terms' magnitudes), not bit-exact; cell
                                      keys, sorted order, neighbour sets and
                                      refinement decisions are unchanged.
                                      Scenes with surface tension or the Morris
                                      term keep the exact sums. Default off */

/* Simulation::Parameters (Simulation.cuh:64-109) + extensions. */
typedef struct asph_parameters {
    asph_domain domain;
    asph_vec3 gravity;
    float speed_of_sound;
    float rest_density;
    float restitution;          /* unused by the reference kernels */
    float viscosity_constant;   /* artificial-viscosity alpha */
    float surface_tension_constant;
    float max_velocity;
    float base_smoothing_radius;
    float base_particle_radius;
    float base_particle_mass;
    uint32_t threads_per_block; /* reference launch width; shapes the split count */
    /* extension: Tait exponent; the reference hard-codes 5 (Common.cuh:8) */
    float tait_gamma;
    /* extension: sloshing as a body force g_eff(t) = g - A sin(2 pi t / T) x^ */
    float slosh_amplitude;
    float slosh_period;
    uint32_t flags;             /* extension: ASPH_FLAG_* */
    /* extension: kinematic viscosity nu (m^2/s) of a Morris laminar term
       (Morris, Fox & Zhu 1997, eq. 6, with mu = rho nu), added after the
       artificial viscosity: a_i += sum_j m_j nu (rho_i + rho_j) / (rho_i rho_j)
       * (r_ij . grad_i W(r, h_j)) / (r^2 + 0.01 h_i^2) * v_ij. 0 = off (the
       reference has no such term). */
    float laminar_viscosity;
} asph_parameters;

/* RefinementParameters::Criterion (RefinementParameters.cuh:82-87) + extension */
#define ASPH_CRITERION_VELOCITY 0u
#define ASPH_CRITERION_INTERFACE 1u
#define ASPH_CRITERION_VORTICITY 2u
#define ASPH_CRITERION_ZONE 3u /* extension: moving spherical refinement zone */

/* SplittingParameters (RefinementParameters.cuh:72-78). */
typedef struct asph_splitting_parameters {
    float epsilon;
    float alpha;
    float center_mass_ratio;
    float vertex_mass_ratio;
} asph_splitting_parameters;

/* {Velocity,Vorticity,Curvature,Interface}Parameters: split then merge threshold. */
typedef struct asph_threshold_pair {
    float split;
    float merge;
} asph_threshold_pair;

/* Extension: spherical zone of radius `radius` whose centre moves on a
   horizontal circle of radius `path_radius` around `center` with period
   `path_period` (0 = static zone). Split inside, merge outside. */
typedef struct asph_zone_parameters {
    asph_vec3 center;
    float radius;
    float path_radius;
    float path_period;
} asph_zone_parameters;

/* RefinementParameters (RefinementParameters.cuh:80-104) + extension. */
typedef struct asph_refinement_parameters {
    uint32_t enabled;
    float min_mass_ratio;
    float max_mass_ratio;
    uint32_t max_particle_count;
    float max_batch_ratio;
    uint32_t initial_cooldown;
    uint32_t cooldown;
    uint32_t criterion_type; /* ASPH_CRITERION_* */
    asph_splitting_parameters splitting;
    asph_threshold_pair velocity;  /* minimalSpeedThreshold, maximalSpeedThreshold */
    asph_threshold_pair vorticity; /* minimal/maximalVorticityThreshold */
    asph_threshold_pair curvature; /* unused by the reference kernels */
    asph_threshold_pair interface_ratio; /* split/merge distanceRatioThreshold */
    asph_zone_parameters zone;     /* extension */
} asph_refinement_parameters;

/* Simulation::DensityInfo (Simulation.cuh:111-120). */
typedef struct asph_density_info {
    float rest_density;
    float min_density;
    float max_density;
    float average_density;
    uint32_t under_density_count;
    uint32_t normal_density_count;
    uint32_t over_density_count;
} asph_density_info;

/* StaticBoundaryDomain::Particle (StaticBoundaryDomain.cuh:17-21). */
typedef struct asph_boundary_particle {
    asph_vec4 position;
    float psi;
} asph_boundary_particle;

/* FluidParticlesData (Simulation.cuh:22-33): device pointers, canonical order
   (index i is the particle the caller created at index i, renumbered only by
   refinement exactly as the reference renumbers). Valid until the next call
   that changes the simulation. Read-only: the library keeps facts about the
   state it owns (a plain simulation's smoothing lengths and masses all equal
   the base values, which its density loop relies on); asph_debug_set_fluid is
   the writer. */
typedef struct asph_fluid_view {
    asph_vec4* positions;
    asph_vec4* velocities;
    asph_vec4* accelerations;
    float* densities;
    float* radii;
    float* smoothing_radii;
    float* masses;
    uint32_t particle_count;
} asph_fluid_view;

/* asph_copy_fluid field mask */
#define ASPH_FIELD_POSITIONS 1u
#define ASPH_FIELD_VELOCITIES 2u
#define ASPH_FIELD_ACCELERATIONS 4u
#define ASPH_FIELD_DENSITIES 8u
#define ASPH_FIELD_RADII 16u
#define ASPH_FIELD_SMOOTHING_RADII 32u
#define ASPH_FIELD_MASSES 64u

typedef struct asph_simulation asph_simulation;

/* Message of the last failed call on this thread ("" if none). */
const char* asph_last_error(void);

/* Fills p / r with the reference's struct defaults (Simulation.cuh:68-70,108;
   RefinementParameters.cuh:8-104) and the extension defaults. */
int asph_default_parameters(asph_parameters* p);
int asph_default_refinement_parameters(asph_refinement_parameters* r);

/* StaticBoundaryDomain::generate (StaticBoundaryDomain.cuh:23-26,
   StaticBoundaryDomain.cu:76-174): wall lattice 0.05 m outside the domain and
   Akinci psi = rest_density / sum W. Call with out == NULL to get the count,
   then with an array of that many records. Host-side, no GPU needed. */
int asph_boundary_generate(const asph_domain* domain, float particle_spacing,
                           float rest_density, float smoothing_radius,
                           asph_boundary_particle* out, uint32_t* count);

/* The same walls with psi computed on the current GPU
   (StaticBoundaryDomain.cu:149-174 replaced by a bucket-grid radius sum
   summed in the host generator's order): identical records, for domains
   that are regenerated while the simulation runs. Synchronous. */
int asph_boundary_generate_device(const asph_domain* domain, float particle_spacing,
                                  float rest_density, float smoothing_radius,
                                  asph_boundary_particle* out, uint32_t* count);

/* No counterpart in the reference. For callers that cannot allocate device
   memory (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc): the records of asph_boundary_generate_device, bit for
   bit, built with no device allocation and no host-device copy, in a fixed
   scratch array of about 54 MiB that the library holds on each GPU it runs on
   until the process ends. out is DEVICE memory of *count records on the
   current GPU (NULL: *count receives the wall count only, no GPU work).
   stream is a hipStream_t (NULL: the null stream); the call returns once the
   work is enqueued, so synchronise the stream before reading out. Limits
   (ASPH_ERR_CAPACITY beyond them): at most ASPH_NOALLOC_MAX_WALLS wall
   particles, ASPH_NOALLOC_MAX_BUCKETS buckets of width 2 smoothing_radius over
   the walls' box and 65,536 values of the three coordinate sequences. Calls
   share their GPU's scratch array: calls that overlap on one GPU (other
   streams or threads) race. */
#define ASPH_NOALLOC_MAX_WALLS 1048576u
#define ASPH_NOALLOC_MAX_BUCKETS 4194304u
int asph_boundary_generate_device_noalloc(const asph_domain* domain, float particle_spacing,
                                          float rest_density, float smoothing_radius,
                                          asph_boundary_particle* out, uint32_t* count, void* stream);

/* createSimulation (Simulation.cuh:138-145, SimulationFactory.cu:14-40).
   The adaptive variant is chosen when refinement != NULL and
   refinement->enabled; it then uses refinement->max_particle_count as the
   fluid capacity, as the reference does (AdaptiveSphSimulation.cu:37).
   positions: host array of count records. boundary: host array of
   boundary_count records (may be NULL when boundary_count == 0). Fluid
   capacities above 2^28 - 64 particles are refused (ASPH_ERR_CAPACITY). */
int asph_create(const asph_parameters* params, const asph_vec4* positions, uint32_t count,
                const asph_boundary_particle* boundary, uint32_t boundary_count,
                const asph_refinement_parameters* refinement,
                uint32_t max_fluid_capacity, uint32_t max_boundary_capacity,
                asph_simulation** out);
int asph_destroy(asph_simulation* sim);

/* Imported particle memory. The reference's caller hands the simulation
   device buffers it owns (the renderer's Vulkan buffers, imported with
   importBuffer, ImportedMemory.cuh:10-26), grouped as
   FluidParticlesDataImportedBuffer and BoundaryParticlesDataImportedBuffer
   (Simulation.cuh:46-59); the simulation keeps the fluid positions,
   velocities, densities and radii in those buffers and writes the boundary
   positions, radii and colours into the boundary ones. Here a buffer is any
   device allocation of the current GPU: asph_buffer_allocate stands in for
   importBuffer (there is no Vulkan interop), and a caller's own hipMalloc'd
   memory serves as well. */
typedef struct asph_device_buffer {
    void* data;  /* device pointer */
    size_t size; /* bytes */
} asph_device_buffer;

/* FluidParticlesDataImportedBuffer: positions and velocities hold vec4
   records, densities and radii floats, each at least the fluid capacity. */
typedef struct asph_fluid_buffers {
    asph_device_buffer positions;
    asph_device_buffer velocities;
    asph_device_buffer densities;
    asph_device_buffer radii;
} asph_fluid_buffers;

/* BoundaryParticlesDataImportedBuffer: positions and colours vec4 records,
   radii floats, each at least the boundary capacity. */
typedef struct asph_boundary_buffers {
    asph_device_buffer positions;
    asph_device_buffer radii;
    asph_device_buffer colors;
} asph_boundary_buffers;

/* importBuffer (ImportedMemory.cuh:22-26) replaced by a device allocation of
   `size` bytes on the current GPU, and its release. */
int asph_buffer_allocate(size_t size, asph_device_buffer* out);
int asph_buffer_release(asph_device_buffer* buffer);
/* Copies between a buffer (from `offset` bytes) and host memory. */
int asph_buffer_download(const asph_device_buffer* buffer, size_t offset, void* host, size_t bytes);
int asph_buffer_upload(const asph_device_buffer* buffer, size_t offset, const void* host, size_t bytes);

/* createSimulation with the reference's full signature (Simulation.cuh:
   138-145): parameters, initial positions, the imported fluid and boundary
   memory, the static boundary domain's particles, the refinement parameters
   (NULL = std::nullopt), and the capacities. The fluid buffers become the
   simulation's own arrays (read and written in place by every update); the
   boundary buffers receive the wall positions, radii (base_particle_radius)
   and colours (0.3, 0.3, 0.3, 0.2) in the domain's order, as
   SphSimulation::initializeBoundaryParticles writes them
   (SphSimulation.cu:287-333), and again on every updateDomain. The buffers
   must outlive the simulation. Other arguments as asph_create. */
int asph_create_imported(const asph_parameters* params, const asph_vec4* positions, uint32_t count,
                         const asph_fluid_buffers* fluid_memory, const asph_boundary_buffers* boundary_memory,
                         const asph_boundary_particle* boundary, uint32_t boundary_count,
                         const asph_refinement_parameters* refinement, uint32_t max_fluid_capacity,
                         uint32_t max_boundary_capacity, asph_simulation** out);

/* BoundaryParticlesData (Simulation.cuh:35-44): device pointers, in the
   boundary domain's order (the simulation's neighbour search keeps its own
   copy sorted by cell). positions, colors and radii are the imported
   boundary buffers of asph_create_imported, or arrays of the simulation;
   psi_values and viscosity_coefficients (domain.friction) are the
   simulation's. Valid until the next updateDomain or asph_destroy. */
typedef struct asph_boundary_view {
    asph_vec4* positions;
    asph_vec4* colors;
    float* psi_values;
    float* viscosity_coefficients;
    float* radii;
    uint32_t particle_count;
} asph_boundary_view;
int asph_get_boundary_view(asph_simulation* sim, asph_boundary_view* out);

/* Simulation::update (Simulation.cuh:126; SphSimulation.cu:101-120,
   AdaptiveSphSimulation.cu:59-85). Synchronous, as in the reference. */
int asph_update(asph_simulation* sim, float dt);

/* Simulation::updateDomain (Simulation.cuh:127-128; SphSimulation.cu:122-134). */
int asph_update_domain(asph_simulation* sim, const asph_domain* domain,
                       const asph_boundary_particle* boundary, uint32_t boundary_count);

/* updateDomain(domain, StaticBoundaryDomain::generate(domain, particle_spacing,
   rest_density, base_smoothing_radius)) as the application calls them
   (App.cpp:143-148 and :377), with the walls, their psi and their binning
   built on the GPU: no particle data passes through the host (for domains
   that move while the simulation runs). Same state as asph_update_domain
   with asph_boundary_generate's records. Synchronous. */
int asph_update_domain_walls(asph_simulation* sim, const asph_domain* domain,
                             float particle_spacing);

/* getFluidParticlesCount / getBoundaryParticlesCount (Simulation.cuh:130-131). */
int asph_fluid_count(asph_simulation* sim, uint32_t* count);
int asph_boundary_count(asph_simulation* sim, uint32_t* count);

/* calculateAverageNeighborCount (Simulation.cuh:133; SphSimulation.cu:237-247). */
int asph_average_neighbor_count(asph_simulation* sim, float* out);

/* getDensityInfo (Simulation.cuh:135; SphSimulation.cu:249-285). */
int asph_get_density_info(asph_simulation* sim, float threshold, asph_density_info* out);

/* enableAdaptiveRefinement (Simulation.cuh:137; AdaptiveSphSimulation.cu:87-90). */
int asph_enable_adaptive_refinement(asph_simulation* sim);

/* Particle-buffer access: the reference hands the renderer the imported
   buffers of FluidParticlesData (SphSimulation.cu:146-161). */
int asph_get_fluid_view(asph_simulation* sim, asph_fluid_view* out);

/* Host download of the fields in `mask` (ASPH_FIELD_*), count records each,
   into the non-NULL destination pointers. */
int asph_copy_fluid(asph_simulation* sim, uint32_t mask, asph_vec4* positions,
                    asph_vec4* velocities, asph_vec4* accelerations, float* densities,
                    float* radii, float* smoothing_radii, float* masses);

#ifdef __cplusplus
}
#endif

#endif /* ASPH_H */
