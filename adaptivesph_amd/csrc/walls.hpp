// Wall particles of StaticBoundaryDomain::generate (StaticBoundaryDomain.cu:
// 76-174): the lattice on the host, psi on the host (boundary.cpp) or on the
// current GPU (walls.hip), both with the same buckets and summation order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/asph.h"

namespace asph
{

struct P4
{
    float x, y, z, w;
};

// Bucket grid of the psi neighbour search: cells of width 2h (FP64) from the
// lowest particle coordinate; a particle's candidates are the particles of
// its 27 surrounding buckets (clamped to the grid).
struct WallBuckets
{
    double lo[3];
    int dims[3];
    double w;
};

// The six faces 0.05 m outside the domain, FP32 accumulated loops, in the
// reference's order.
std::vector<P4> wallLattice(const asph_domain& d, float spacing);
WallBuckets wallBuckets(const std::vector<P4>& pts, float h);

// psi[i] = rho0 * (1 / sum W(|p_i - p_j|, h)) over the particles j within the
// 2h support (strict d^2 < (2h)^2, i itself included), summed in ascending j.
void wallPsi(const std::vector<P4>& pts, float h, float rho0, std::vector<float>& psi);
void wallPsiDevice(const std::vector<P4>& pts, float h, float rho0, std::vector<float>& psi);

// The lattice as six product faces over the FP32 coordinate sequences of
// wallLattice (x, y and z from the min corner 0.05 m outside the domain in
// steps of `spacing`): face f holds, in order, for each a of its outer
// sequence and each b of its inner one, the particle with coordinate a on
// axisA, b on axisB and `fixed` on the remaining axis. coords = X ++ Y ++ Z.
struct WallFace
{
    uint32_t offA, nA, offB, nB;  // ranges of coords
    int axisA, axisB, axisFixed;
    float fixed;
    uint32_t start;  // first particle of the face
};

struct WallFaces
{
    std::vector<float> coords;
    WallFace face[6];
    uint32_t count;
    WallBuckets buckets;  // wallBuckets of the lattice, for psi with smoothing radius h
};

WallFaces wallFaces(const asph_domain& d, float spacing, float h);

// The three coordinate sequences of wallLattice, counted without storing
// them: min corner mn, loop bound mx, length, second and last value per axis
// (the sequences increase). Fails when a sequence does not advance
// (spacing below half an ulp of a coordinate) or the lengths together exceed
// maxTotal (kWallSeq* codes).
struct WallSequences
{
    float mn[3], mx[3];
    uint32_t len[3];
    float second[3], last[3];
};
constexpr int kWallSeqOk = 0;
constexpr int kWallSeqStalls = 1;   // a sequence does not advance
constexpr int kWallSeqTooLong = 2;  // more than maxTotal values
int wallSequences(const asph_domain& d, float spacing, uint64_t maxTotal, WallSequences* out);
// Face table, particle count and bucket grid of wallFaces from the sequences
// (coords of wf left untouched).
void wallFaceTable(const WallSequences& sq, float h, WallFaces* wf);

// StaticBoundaryDomain::generate with the lattice, psi and output records all
// built on the current GPU in the library's fixed scratch array (no device
// allocation), enqueued on stream s; out is device memory of *count records
// (NULL: *count receives the wall count only).
int boundaryGenerateNoAlloc(const asph_domain& d, float spacing, float rho0, float h, asph_boundary_particle* out,
                            uint32_t* count, hipStream_t s);

// psi of n device-resident wall positions (w ignored) into psi[n] (device);
// synchronous on stream s.
struct F4;
void wallPsiOnDevice(const F4* pos, uint32_t n, const WallBuckets& wb, float h, float rho0, float* psi,
                     hipStream_t s);

// Device destinations of a scene's boundary particles (the engine's pools,
// unoffset: the scene's bOff and cellBase are applied by wallsOnDevice).
struct SceneDev;
struct WallTargets
{
    F4* bs0;
    float* bvisc;
    uint32_t* bcellOff;
    F4* bb0;
    F4* bb1;
    F4* bg0;
    F4* bg1;
    float* bnear2;
    F4* lattice = nullptr;        // optional: positions in the generator's order
    float* latticePsi = nullptr;  // optional: their psi
};

// The walls of the domain of scene description d (whose grid, pool offsets,
// friction, h0 and rest density are already those of the new domain), built
// on the current GPU straight into the engine's boundary arrays, binned as
// the host ingestion bins them; returns the end of the slot-0 omission and
// whether a wall particle lies in an interior cell. wf.count > 0. Synchronous.
void wallsOnDevice(const WallFaces& wf, const SceneDev& d, const WallTargets& out, uint32_t* lostEnd, bool* inside);

// asph_boundary_generate (onDevice false) and asph_boundary_generate_device.
int boundaryGenerate(const asph_domain& d, float spacing, float rho0, float h, asph_boundary_particle* out,
                     uint32_t* count, bool onDevice);

}  // namespace asph
