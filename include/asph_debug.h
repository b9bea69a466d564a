/*
 * asph_debug.h - test hooks of the C ABI. None of these has a counterpart in
 * the reference: they expose intermediate products of the last update (the
 * neighbour grid, neighbour sets, refinement decisions) and let a test load an
 * arbitrary particle state, so parity tests can compare each stage with the
 * oracle on identical FP32 inputs.
 */
#ifndef ASPH_DEBUG_H
#define ASPH_DEBUG_H

#include "asph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replace the fluid state (count records per array, count <= capacity).
   Arrays are host pointers; any may be NULL to keep that field. */
int asph_debug_set_fluid(asph_simulation* sim, uint32_t count, const asph_vec4* positions,
                         const asph_vec4* velocities, const asph_vec4* accelerations,
                         const float* densities, const float* radii,
                         const float* smoothing_radii, const float* masses);

/* Set the refinement frame counter (AdaptiveSphSimulation.cuh:43). */
int asph_debug_set_frame(asph_simulation* sim, uint32_t frame);

/* Grid dimensions (gx, gy, gz) of the fluid neighbour grid. */
int asph_debug_grid_dims(asph_simulation* sim, int32_t dims[3]);

/* Fluid neighbour grid of the last update in the reference's format
   (NeighborGrid.cuh:18-26): sorted cell keys and particle indices (count
   records each) and inclusive cell start/end (gx*gy*gz records each, -1 for
   empty), including the reference's slot-0 omission unless
   ASPH_FLAG_FIX_CELL_START is set. Any pointer may be NULL. */
int asph_debug_grid(asph_simulation* sim, int32_t* sorted_keys, int32_t* sorted_indices,
                    int32_t* cell_start, int32_t* cell_end);

/* Per-particle fluid neighbour sets of the last update, as counts (j != i,
   d^2 <= (2 h_j)^2 over the reference's candidate cells) and an
   order-independent 64-bit fingerprint of the set of neighbour indices. */
int asph_debug_neighbor_sets(asph_simulation* sim, uint32_t* counts, uint64_t* fingerprints);

/* Decisions of the last refinement (AdaptiveSphSimulation.cu:102-252). */
typedef struct asph_refinement_trace {
    uint32_t refined;          /* 1 if a refinement ran in the last update */
    uint32_t count_before;     /* particles before merging */
    uint32_t merge_eligible;   /* length of the merge list */
    uint32_t merge_pairs;      /* number of merged pairs */
    uint32_t count_after_merge;
    uint32_t split_selected;   /* length of the split list (k) */
    uint32_t split_performed;  /* parents actually split */
    uint32_t count_after;      /* particles after splitting */
} asph_refinement_trace;

/* trace: summary. The arrays receive, when non-NULL: the merge list
   (merge_eligible ids), the merge candidate of each particle (count_before
   records; noCandidate = 0xffffffff, unlisted particles 0 as in the
   reference), the keep ids of merged pairs in ascending order (merge_pairs
   records) and the parents split (split_performed ids, in split order). */
int asph_debug_refinement(asph_simulation* sim, asph_refinement_trace* trace,
                          uint32_t* merge_list, uint32_t* merge_candidates,
                          uint32_t* pair_keep_ids, uint32_t* split_parents);

/* Checks, on `count` pseudo-random operand pairs (seeded) with magnitudes in
   [2^-40, 2^40], that the kernels' division through a shared reciprocal gives
   the same bits as the FP32 division, and on `count` pseudo-random values
   (+0 or [2^-96, 2^127)) that their range-restricted square root gives the
   bits of sqrtf; *mismatches receives the number of results that differ
   (0 expected). Uses the current device; synchronous. */
int asph_debug_division_check(uint64_t seed, uint64_t count, uint64_t* mismatches);

#ifdef __cplusplus
}
#endif

#endif /* ASPH_DEBUG_H */
