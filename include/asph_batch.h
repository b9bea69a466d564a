/*
 * asph_batch.h - batched and served stepping on one GPU. No counterpart in
 * the reference (which steps one scene per process): many independent scenes
 * share one set of device pools and are advanced together, each kernel of a
 * step covering every scene (segmented by a scene table), optionally replayed
 * as one captured hipGraph. Scenes may be added and removed between steps
 * (continuous batching); their state stays resident in device memory.
 */
#ifndef ASPH_BATCH_H
#define ASPH_BATCH_H

#include "asph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* asph_batch_config.flags */
#define ASPH_BATCH_GRAPH 1u    /* capture one step as a hipGraph and replay it */
#define ASPH_BATCH_ADAPTIVE 2u /* allocate refinement scratch (adaptive scenes) */

typedef struct asph_batch_config {
    uint32_t max_particles; /* fluid particle pool (sum of scene capacities), at most 2^28 - 64 */
    uint32_t max_boundary;  /* boundary particle pool */
    uint32_t max_cells;     /* grid cell pool (sum over scenes of cells + 1) */
    uint32_t max_scenes;    /* scene table size */
    uint32_t list_capacity; /* neighbour-list entries per particle (0 = 128, or 640
                               with ASPH_BATCH_ADAPTIVE) */
    uint32_t flags;         /* ASPH_BATCH_* */
} asph_batch_config;

typedef struct asph_batch asph_batch;

int asph_batch_create(const asph_batch_config* config, asph_batch** out);
int asph_batch_destroy(asph_batch* batch);

/* Adds a scene (same meaning of arguments as asph_create; capacity is the
   fluid capacity of a non-adaptive scene) stepped with time step dt. */
int asph_batch_add_scene(asph_batch* batch, const asph_parameters* params, const asph_vec4* positions,
                         uint32_t count, const asph_boundary_particle* boundary, uint32_t boundary_count,
                         const asph_refinement_parameters* refinement, uint32_t capacity, float dt,
                         uint32_t* scene_id);
int asph_batch_remove_scene(asph_batch* batch, uint32_t scene_id);

/* asph_update_domain_walls for one scene of the batch (between steps): the
   scene's box moves or resizes and its walls are rebuilt on the GPU; the
   other scenes are untouched (Simulation::updateDomain, SphSimulation.cu:
   122-134, per scene). */
int asph_batch_update_domain_walls(asph_batch* batch, uint32_t scene_id, const asph_domain* domain,
                                   float particle_spacing);

/* Enqueue `steps` steps of every scene; returns without waiting. */
int asph_batch_step(asph_batch* batch, uint32_t steps);
int asph_batch_synchronize(asph_batch* batch);

/* One step timed with HIP events on the batch's stream (synchronous). */
int asph_batch_step_timed(asph_batch* batch, float* milliseconds);

int asph_batch_scene_count(asph_batch* batch, uint32_t scene_id, uint32_t* count);

/* calculateAverageNeighborCount (Simulation.cuh:133; SphSimulation.cu:237-247)
   of one scene of the batch, over the grid of the most recent step. */
int asph_batch_average_neighbor_count(asph_batch* batch, uint32_t scene_id, float* out);
int asph_batch_copy_scene(asph_batch* batch, uint32_t scene_id, uint32_t mask, asph_vec4* positions,
                          asph_vec4* velocities, asph_vec4* accelerations, float* densities, float* radii,
                          float* smoothing_radii, float* masses);

/* Launch shape: threads per block of the streaming kernels (64..1024, a
   multiple of 64) and of the neighbour kernels (64, 128 or 256), radix digit
   width of the cell-key sort (4, 6, 7, 8 or 10 bits; 0 picks the width with
   the fewest passes over the active key width) and graph replay on/off.
   Results do not depend on the launch shape. */
int asph_batch_set_launch(asph_batch* batch, uint32_t block_stream, uint32_t block_neighbor, uint32_t radix_bits,
                          uint32_t use_graph);

/* Schedule of the neighbour kernels. query_tile orders each wave's particles:
   0 in sorted (cell) order; 1 or 2 within each cell of more than 64 particles
   by sub-cell tiles of 2x2x2 or 4x4x4 (Morton order); 3 by 4x4x4 sub-cells
   in x-major order (slabs of the cell); 4 per scene, the sorted order until
   the scene's waves loosen (their candidate filters keep more than 8.5
   survivors per neighbour of a particle, as the flow mixes each cell's
   particles), then 3; -1 the default (4). force_blocks_per_cu: 0 launches the force
   kernels with one thread per particle, 1..64 as a persistent grid of that
   many blocks per compute unit. Results do not depend on the schedule. */
int asph_batch_set_schedule(asph_batch* batch, int32_t query_tile, uint32_t force_blocks_per_cu);

/* Particles of all scenes stepped by the most recent step (their count at
   the start of that step, before refinement). Synchronous. */
int asph_batch_step_particles(asph_batch* batch, uint32_t* count);

/* Device bytes allocated by the batch and particle slots in use. */
int asph_batch_memory(asph_batch* batch, uint64_t* allocated, uint64_t* particles_used);

/* One eager step with a HIP event between stages; ms[i] receives the time of
   stage i (up to capacity entries), *count the number of stages. */
int asph_batch_profile_step(asph_batch* batch, float* ms, uint32_t capacity, uint32_t* count);
/* The same eager step, enqueued without waiting (between graph-replayed steps
   of a timed run); asph_batch_profile_collect waits for the batch's stream
   and returns each stage's time averaged over the steps so enqueued since the
   last collection (*samples of them). */
int asph_batch_profile_step_async(asph_batch* batch);
int asph_batch_profile_collect(asph_batch* batch, float* ms, uint32_t capacity, uint32_t* count, uint32_t* samples);
const char* asph_profile_stage_name(uint32_t stage);

/* Device bytes one scene of this shape occupies inside a batch
   (list_capacity 0: the batch default for `adaptive`). */
uint64_t asph_scene_bytes(uint32_t capacity, uint32_t boundary_count, uint32_t cells, uint32_t adaptive,
                          uint32_t list_capacity);

/* Cell count of the neighbour grid of a domain for smoothing radius h. */
int asph_grid_cells(const asph_domain* domain, float smoothing_radius, uint32_t* cells);

/* Device name and compute-unit count as text. */
int asph_device_info(char* buffer, uint32_t length);

#ifdef __cplusplus
}
#endif

#endif /* ASPH_BATCH_H */
