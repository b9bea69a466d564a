// Launchers of the HIP kernels (sort.hip, sph.hip). All launches are
// asynchronous on the given stream and read element counts from device memory
// where noted, so one step is a static sequence of launches (graph-capturable).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device_common.hpp"

namespace asph
{

// ------------------------------------------------------------ sort.hip --
size_t radixHistWords(uint32_t nMax, int radixBits);
// Digit width of the cell-key sort: `requested` (4, 6, 7, 8 or 10), or for 0
// the width with the fewest passes over keyBits bits.
int radixDigitBits(int requested, int keyBits);
constexpr int kMaxRadixBits = 10;
size_t scanTileWords(uint32_t nMax);
// Returns 0 when the result is in (keysA, valsA), 1 when in (keysB, valsB).
int radixSortU32(uint32_t* keysA, uint32_t* valsA, uint32_t* keysB, uint32_t* valsB, const uint32_t* nPtr,
                 uint32_t nMax, int bits, int radixBits, uint32_t* hist, uint32_t* digitTotals, hipStream_t stream);
int radixSortU64(uint64_t* keysA, uint32_t* valsA, uint64_t* keysB, uint32_t* valsB, const uint32_t* nPtr,
                 uint32_t nMax, int bits, uint32_t* hist, uint32_t* digitTotals, hipStream_t stream);
// out[i] = sum in[0, i) for i < n (n = *nPtr if nPtr else nMax); *total = sum.
void exclusiveScanU32(uint32_t* in, uint32_t* out, const uint32_t* nPtr, uint32_t nMax, uint32_t* tileSums,
                      uint32_t* total, bool zeroInput, hipStream_t stream);

// ------------------------------------------------------------- sph.hip --

// Device pointers of the engine's pools.
struct Ctx
{
    // scene tables
    const SceneDev* scenes;
    SceneState* states;
    uint32_t sceneSlots;        // number of scene slots (table size)
    uint32_t* denseStarts;      // [sceneSlots] exclusive prefix of counts
    uint32_t* sortedStarts;     // [sceneSlots] first sorted slot, in cell-base order
    const uint32_t* cellOrder;  // [sceneSlots] scene ids in cell-base order
    const uint32_t* cellBaseOrder;  // [sceneSlots] their first cells
    const uint32_t* nOrderPtr;  // device: active scenes in cellOrder
    uint32_t* nTotal;           // device: active particles this step

    // canonical particle arrays (interface layout)
    F4* pos;
    F4* vel;
    F4* acc;
    float* rho;
    float* rad;
    float* h;
    float* m;
    uint32_t pCap;  // pool capacity (particles); stride of the neighbour list rows
    // Launch extents of the active scenes (host-set when scenes are added or
    // removed): nMax bounds the particles of a step (sum of the active
    // scenes' capacities), cExtent the cells in use (highest cell slot + 1).
    uint32_t nMax;
    uint32_t cExtent;
    // force loops present among the active scenes (host-set): kForceExact
    // (kForce) and kForceTolerance (kForceTol)
    uint32_t forceKinds;

    // neighbour grid
    uint32_t* keysA;
    uint32_t* valsA;
    uint32_t* keysB;
    uint32_t* valsB;
    uint32_t* sortedKeys;  // result buffers of the sort (set per step)
    uint32_t* sortedVals;
    uint32_t* cellOff;    // [cCap + 1]
    uint32_t cCap;

    // sorted shadow copies
    F4* s0;      // (x, y, z, (2h)^2)
    F4* s1;      // (vx, vy, vz, m)
    F4* shm;     // (h, m, kWCoef / h^3, kWDCoef / h^4)
    F4* fa;      // (x, y, z, p / rho^2), written with the density
    F4* fb;      // (vx, vy, vz, rho), written with the density
    uint32_t* sscene;
    uint32_t* ncount;
    // particles whose neighbours overflow the list (sorted slots), and their
    // number (reset by kPrep each step)
    uint32_t* overflowList;
    uint32_t* overflowCount;
    // particles whose boundary forces are added by kBoundaryForce (sorted
    // slots; walls within reach), and their number (reset by kPrep each step)
    uint32_t* wallQueue;
    uint32_t* wallQueueCount;
    uint32_t* queryAny;  // device: some scene takes the query order this step (kPrep)
    // list builder's wave order (heavy waves first): per wave of 64 query
    // positions its cost in the last step (survivors, or kCostPerLane) and
    // whether it is listed; the list; sampled cost statistics
    // (sum, waves, threshold)
    uint32_t* wcost;
    uint32_t* wheavy;
    uint32_t* heavyList;
    uint32_t* heavyCount;
    uint32_t* wstat;
    uint32_t* list;  // [listCap * pCap rounded up to 64], 4-entry chunks, wave-blocked: entry k of slot t
                     // at (((t / 64) * (listCap / 4) + k / 4) * 64 + t % 64) * 4 + k % 4
    uint32_t listCap;  // a multiple of 4
    // per aligned group of kGroup sorted slots: bounding box and max (2h)^2
    F4* gbox0;  // (min x, min y, min z, max (2h)^2)
    F4* gbox1;  // (max x, max y, max z, 0)

    // boundary (sorted by cell per scene)
    F4* bs0;  // (x, y, z, psi)
    float* bvisc;
    uint32_t* bcellOff;  // [cCap]: absolute boundary slot of each cell start
    // [cCap]: bounding box of the boundary particles keyed to each cell
    // (min xyz, max xyz; empty cells hold an inverted box)
    F4* bbox0;
    F4* bbox1;
    F4* bgbox0;  // at the first slot of each chunk of kBoundaryGroup slots of a
    F4* bgbox1;  // cell's boundary range: the box of the chunk's positions
    // per cell: a lower bound of the squared distance from any point of the
    // cell to any boundary particle (capped at two cell widths)
    float* bnear2;

    // refinement scratch (dense-indexed)
    uint64_t* ckeysA;
    uint32_t* cvalsA;
    uint64_t* ckeysB;
    uint32_t* cvalsB;
    uint64_t* csortedKeys;
    uint32_t* csortedVals;
    uint32_t* cand;
    uint32_t* slotOf;  // [pCap] sorted slot of each canonical particle (written by kGather; adaptive engines)
    // device: criterion keys of the current refinement phase (merge: set by
    // kPrep; split: by kAfterMerge); the participating scenes' keys are laid
    // out back to back from each scene's segStart
    uint32_t* critCount;
    // Query order of the list builder and the force kernels (null: the
    // sorted order; LaunchCfg::queryTile): query position p handles sorted
    // slot qperm[p], and slot t's neighbour list sits at query position
    // qpos[t]. Within each cell the slots are grouped by sub-cell, so the 64
    // queries of a wave are near each other even where refinement packs a
    // cell with daughters in index order.
    uint32_t* qperm;
    uint32_t* qpos;
    uint32_t queryTileBits;  // query order mode (with qperm): 1, 2, 3 or kQueryAuto
    uint32_t* removal;
    uint32_t* keep;
    uint32_t* rprefix;
    F4* xpos;
    F4* xvel;
    F4* xacc;
    float* xrho;
    float* xrad;
    float* xh;
    float* xm;
    uint32_t* mergeListDbg;  // optional capture of the merge list (sorted vals)
};

struct LaunchCfg
{
    uint32_t blockStream = 256;  // streaming kernels
    uint32_t blockNeighbor = 256;  // neighbour kernels (64, 128 or 256)
    int radixBits = 0;  // 0: chosen from the key width (radixDigitBits)
    // Query order of the neighbour kernels (kQueryOrder): 0 the sorted (cell)
    // order; k = 1 or 2: within each cell of more than 64 particles, tiles of
    // 2^k x 2^k x 2^k sub-cells in Morton order; -1: 2 in adaptive engines,
    // 0 in plain ones.
    int queryTile = -1;
    // Grid of the force kernels: 0 one thread per query position; otherwise a
    // persistent grid of this many blocks per compute unit (grid-stride loop)
    uint32_t forceBlocksPerCU = 0;
    uint32_t forceGrid = 0;  // set by the engine: forceBlocksPerCU x compute units
};

constexpr uint32_t kForceExact = 1U;
constexpr uint32_t kForceTolerance = 2U;

// Step stages, in call order.
void launchPrep(const Ctx& c, hipStream_t s);
void launchKickDriftKey(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchGather(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchQueryOrder(const Ctx& c, hipStream_t s);
void launchNeighbors(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchDensity(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchForce(const Ctx& c, const LaunchCfg& lc, hipStream_t s);

// Refinement stages (AdaptiveSphSimulation::performAdaptiveRefinement).
void launchMergeCriteria(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchMergeSelect(const Ctx& c, hipStream_t s);
void launchMergeCandidates(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchResolvePairs(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchMergeApply(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchCompact(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchAfterMerge(const Ctx& c, hipStream_t s);
void launchCopyBack(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchSplitCriteria(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchSplitSelect(const Ctx& c, hipStream_t s);
void launchSplitApply(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchKickRefined(const Ctx& c, const LaunchCfg& lc, hipStream_t s);
void launchCopyMergeList(const Ctx& c, hipStream_t s);

// Diagnostics of one scene (calculateAverageNeighborCount, getDensityInfo,
// neighbour-set fingerprints).
void launchNeighborSets(const Ctx& c, uint32_t scene, uint32_t count, uint32_t* counts, uint64_t* prints,
                        uint32_t* total, hipStream_t s);
void launchDensityInfo(const Ctx& c, uint32_t scene, uint32_t count, float threshold, double* partials,
                       uint32_t* upartials, hipStream_t s);

// Adds to *bad the pairs among `count` pseudo-random ones where the
// shared-reciprocal division differs from a / b (0 expected).
void launchDivisionCheck(uint64_t seed, uint64_t count, unsigned long long* bad, hipStream_t s);
constexpr uint32_t kInfoBlocks = 256;

}  // namespace asph
