// Wall particles of StaticBoundaryDomain::generate (StaticBoundaryDomain.cu:
// 76-174): the lattice on the host, psi on the host (boundary.cpp) or on the
// current GPU (walls.hip), both with the same buckets and summation order.
#pragma once

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

// asph_boundary_generate (onDevice false) and asph_boundary_generate_device.
int boundaryGenerate(const asph_domain& d, float spacing, float rho0, float h, asph_boundary_particle* out,
                     uint32_t* count, bool onDevice);

}  // namespace asph
