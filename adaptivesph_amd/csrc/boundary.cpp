// StaticBoundaryDomain::generate (StaticBoundaryDomain.cu:76-174) on the host:
// the wall lattice 0.05 m outside the domain (six faces, FP32 accumulated
// loops, in the reference's order) and Akinci psi = rest_density / delta with
// delta = sum over wall particles within the 2h support (strict d^2 < (2h)^2,
// the nanoflann radius test) of W(d, h), including the particle itself. The
// reference uses a kd-tree; this uses a uniform bucket grid and sums the hits
// in ascending particle index.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/asph.h"
#include "walls.hpp"

namespace asph
{

namespace
{

float wendlandHost(float d, float h)
{
    constexpr float pi = 3.141592653589793F;
    constexpr float coef = 21.F / (16.F * pi);
    const float q = d / h;
    if (q > 2.0F)
    {
        return 0.0F;
    }
    const float h3 = (h * h) * h;
    const float tmp = 1.0F - 0.5F * q;
    const float tmp4 = (tmp * tmp) * (tmp * tmp);
    return ((coef / h3) * tmp4) * (2.0F * q + 1.0F);
}

}  // namespace

std::vector<P4> wallLattice(const asph_domain& d, float spacing)
{
    std::vector<P4> out;
    const float off = 0.05F;
    const float mnx = d.min.x - off, mny = d.min.y - off, mnz = d.min.z - off;
    const float mxx = d.max.x + off, mxy = d.max.y + off, mxz = d.max.z + off;
    for (float x = mnx; x < mxx; x += spacing)
    {
        for (float z = mnz; z < mxz; z += spacing)
        {
            out.push_back({x, mny, z, 0.F});
        }
    }
    for (float x = mnx; x < mxx; x += spacing)
    {
        for (float z = mnz; z < mxz; z += spacing)
        {
            out.push_back({x, mxy, z, 0.F});
        }
    }
    for (float y = mny + spacing; y < mxy; y += spacing)
    {
        for (float z = mnz; z < mxz; z += spacing)
        {
            out.push_back({mnx, y, z, 0.F});
        }
    }
    for (float y = mny + spacing; y < mxy; y += spacing)
    {
        for (float z = mnz; z < mxz; z += spacing)
        {
            out.push_back({mxx, y, z, 0.F});
        }
    }
    for (float x = mnx + spacing; x < mxx; x += spacing)
    {
        for (float y = mny + spacing; y < mxy; y += spacing)
        {
            out.push_back({x, y, mnz, 0.F});
        }
    }
    for (float x = mnx + spacing; x < mxx; x += spacing)
    {
        for (float y = mny + spacing; y < mxy; y += spacing)
        {
            out.push_back({x, y, mxz, 0.F});
        }
    }
    return out;
}

int wallSequences(const asph_domain& d, float spacing, uint64_t maxTotal, WallSequences* out)
{
    // the loops of wallLattice: X, Y, Z from the min corner, counted without
    // storing them (the values are the same FP32 additions wherever they run)
    const float off = 0.05F;
    const float mn[3] = {d.min.x - off, d.min.y - off, d.min.z - off};
    const float mx[3] = {d.max.x + off, d.max.y + off, d.max.z + off};
    uint64_t total = 0;
    for (int a = 0; a < 3; ++a)
    {
        out->mn[a] = mn[a];
        out->mx[a] = mx[a];
        uint32_t len = 0;
        float second = mn[a];
        float last = mn[a];
        for (float v = mn[a]; v < mx[a];)
        {
            if (len == 1U)
            {
                second = v;
            }
            last = v;
            ++len;
            if (++total > maxTotal)
            {
                return kWallSeqTooLong;
            }
            const float next = v + spacing;
            if (!(next > v))
            {
                return kWallSeqStalls;  // spacing below half an ulp of v (or not a positive number)
            }
            v = next;
        }
        out->len[a] = len;
        out->second[a] = second;
        out->last[a] = last;
    }
    return kWallSeqOk;
}

void wallFaceTable(const WallSequences& sq, float h, WallFaces* wf)
{
    // coordinate sequences back to back: X at 0, then Y, then Z; the faces
    // that start one step in begin at the sequence's second value
    const uint32_t seqOff[3] = {0U, sq.len[0], sq.len[0] + sq.len[1]};
    struct Range
    {
        uint32_t off, n;
        float lo, hi;  // smallest and largest value of the range (the sequences increase)
    };
    auto full = [&](int a) { return Range {seqOff[a], sq.len[a], sq.mn[a], sq.last[a]}; };
    auto tail = [&](int a) {
        return sq.len[a] == 0U ? Range {seqOff[a], 0U, 0.F, 0.F}
                               : Range {seqOff[a] + 1U, sq.len[a] - 1U, sq.second[a], sq.last[a]};
    };
    struct Spec
    {
        Range A, B;
        int axisA, axisB, axisFixed;
        float fixed;
    };
    const float* mn = sq.mn;
    const float* mx = sq.mx;
    const Spec spec[6] = {
        {full(0), full(2), 0, 2, 1, mn[1]}, {full(0), full(2), 0, 2, 1, mx[1]},
        {tail(1), full(2), 1, 2, 0, mn[0]}, {tail(1), full(2), 1, 2, 0, mx[0]},
        {tail(0), tail(1), 0, 1, 2, mn[2]}, {tail(0), tail(1), 0, 1, 2, mx[2]},
    };
    uint64_t start = 0;
    double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX};
    double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int f = 0; f < 6; ++f)
    {
        const Spec& sp = spec[f];
        WallFace& w = wf->face[f];
        w.offA = sp.A.off;
        w.nA = sp.A.n;
        w.offB = sp.B.off;
        w.nB = sp.B.n;
        w.axisA = sp.axisA;
        w.axisB = sp.axisB;
        w.axisFixed = sp.axisFixed;
        w.fixed = sp.fixed;
        w.start = static_cast<uint32_t>(std::min<uint64_t>(start, UINT32_MAX));
        const uint64_t n = static_cast<uint64_t>(w.nA) * w.nB;
        start += n;
        if (n == 0)
        {
            continue;
        }
        // the face's extent (the bucket grid spans the particles themselves)
        lo[w.axisA] = std::min(lo[w.axisA], static_cast<double>(sp.A.lo));
        hi[w.axisA] = std::max(hi[w.axisA], static_cast<double>(sp.A.hi));
        lo[w.axisB] = std::min(lo[w.axisB], static_cast<double>(sp.B.lo));
        hi[w.axisB] = std::max(hi[w.axisB], static_cast<double>(sp.B.hi));
        lo[w.axisFixed] = std::min(lo[w.axisFixed], static_cast<double>(w.fixed));
        hi[w.axisFixed] = std::max(hi[w.axisFixed], static_cast<double>(w.fixed));
    }
    if (start >= (1ULL << 31))
    {
        throw std::runtime_error("too many wall particles");
    }
    wf->count = static_cast<uint32_t>(start);
    wf->buckets.w = static_cast<double>(2.0F * h);
    for (int a = 0; a < 3; ++a)
    {
        wf->buckets.lo[a] = wf->count == 0 ? 0.0 : lo[a];
        wf->buckets.dims[a] =
            wf->count == 0 ? 1 : std::max(1, static_cast<int>(std::floor((hi[a] - lo[a]) / wf->buckets.w)) + 1);
    }
}

WallFaces wallFaces(const asph_domain& d, float spacing, float h)
{
    WallSequences sq;
    if (wallSequences(d, spacing, 1ULL << 31, &sq) != kWallSeqOk)
    {
        throw std::runtime_error("wall coordinate sequences do not advance or are too long");
    }
    WallFaces wf;
    wf.coords.reserve(static_cast<size_t>(sq.len[0]) + sq.len[1] + sq.len[2]);
    for (int a = 0; a < 3; ++a)
    {
        float v = sq.mn[a];
        for (uint32_t k = 0; k < sq.len[a]; ++k)
        {
            wf.coords.push_back(v);
            v = v + spacing;
        }
    }
    wallFaceTable(sq, h, &wf);
    return wf;
}

WallBuckets wallBuckets(const std::vector<P4>& pts, float h)
{
    WallBuckets wb;
    double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int a = 0; a < 3; ++a)
    {
        wb.lo[a] = DBL_MAX;
    }
    for (const auto& q : pts)
    {
        const double v[3] = {q.x, q.y, q.z};
        for (int a = 0; a < 3; ++a)
        {
            wb.lo[a] = std::min(wb.lo[a], v[a]);
            hi[a] = std::max(hi[a], v[a]);
        }
    }
    wb.w = static_cast<double>(2.0F * h);
    for (int a = 0; a < 3; ++a)
    {
        wb.dims[a] = std::max(1, static_cast<int>(std::floor((hi[a] - wb.lo[a]) / wb.w)) + 1);
    }
    return wb;
}

void wallPsi(const std::vector<P4>& pts, float h, float rho0, std::vector<float>& psi)
{
    const size_t n = pts.size();
    psi.assign(n, 0.F);
    if (n == 0)
    {
        return;
    }
    const float R = 2.0F * h;
    const float R2 = R * R;
    const WallBuckets wb = wallBuckets(pts, h);
    const double* lo = wb.lo;
    const int* dims = wb.dims;
    const double w = wb.w;
    auto bucket = [&](const P4& q, int a) {
        const double v = a == 0 ? q.x : (a == 1 ? q.y : q.z);
        const int b = static_cast<int>(std::floor((v - lo[a]) / w));
        return std::min(std::max(b, 0), dims[a] - 1);
    };
    const size_t nb = static_cast<size_t>(dims[0]) * dims[1] * dims[2];
    std::vector<uint32_t> start(nb + 1, 0);
    std::vector<uint32_t> which(n);
    for (size_t i = 0; i < n; ++i)
    {
        const size_t b = bucket(pts[i], 0) + dims[0] * (bucket(pts[i], 1) + static_cast<size_t>(dims[1]) * bucket(pts[i], 2));
        which[i] = static_cast<uint32_t>(b);
        start[b + 1]++;
    }
    for (size_t b = 0; b < nb; ++b)
    {
        start[b + 1] += start[b];
    }
    std::vector<uint32_t> items(n);
    {
        std::vector<uint32_t> fill(start.begin(), start.end() - 1);
        for (size_t i = 0; i < n; ++i)
        {
            items[fill[which[i]]++] = static_cast<uint32_t>(i);
        }
    }
    auto dist2 = [](const P4& q, const P4& o) {
        float d2 = 0.F;
        const float dx = q.x - o.x;
        d2 = d2 + dx * dx;
        const float dy = q.y - o.y;
        d2 = d2 + dy * dy;
        const float dz = q.z - o.z;
        d2 = d2 + dz * dz;
        return d2;
    };
    auto work = [&](size_t begin, size_t end) {
        std::vector<uint32_t> hits;
        for (size_t i = begin; i < end; ++i)
        {
            const P4& q = pts[i];
            const int bx = bucket(q, 0), by = bucket(q, 1), bz = bucket(q, 2);
            hits.clear();
            for (int z = std::max(bz - 1, 0); z <= std::min(bz + 1, dims[2] - 1); ++z)
            {
                for (int y = std::max(by - 1, 0); y <= std::min(by + 1, dims[1] - 1); ++y)
                {
                    for (int x = std::max(bx - 1, 0); x <= std::min(bx + 1, dims[0] - 1); ++x)
                    {
                        const size_t b = x + dims[0] * (y + static_cast<size_t>(dims[1]) * z);
                        for (uint32_t s = start[b]; s < start[b + 1]; ++s)
                        {
                            if (dist2(q, pts[items[s]]) < R2)
                            {
                                hits.push_back(items[s]);
                            }
                        }
                    }
                }
            }
            std::sort(hits.begin(), hits.end());
            float sum = 0.F;
            for (const uint32_t j : hits)
            {
                sum = sum + wendlandHost(std::sqrt(dist2(q, pts[j])), h);
            }
            const float volume = 1.0F / sum;
            psi[i] = rho0 * volume;
        }
    };
    const unsigned threads = std::max(1U, std::min(16U, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    const size_t chunk = (n + threads - 1) / threads;
    for (unsigned t = 0; t < threads; ++t)
    {
        const size_t b = t * chunk;
        const size_t e = std::min(n, b + chunk);
        if (b < e)
        {
            pool.emplace_back(work, b, e);
        }
    }
    for (auto& th : pool)
    {
        th.join();
    }
}

int boundaryGenerate(const asph_domain& d, float spacing, float rho0, float h, asph_boundary_particle* out,
                     uint32_t* count, bool onDevice)
{
    WallSequences sq;
    if (wallSequences(d, spacing, 1ULL << 31, &sq) != kWallSeqOk)
    {
        throw std::runtime_error("wall coordinate sequences do not advance (spacing too small for the domain's "
                                 "coordinates) or are too long");
    }
    const std::vector<P4> pts = wallLattice(d, spacing);
    if (out == nullptr)
    {
        *count = static_cast<uint32_t>(pts.size());
        return ASPH_OK;
    }
    if (*count < pts.size())
    {
        return ASPH_ERR_CAPACITY;
    }
    std::vector<float> psi;
    if (onDevice)
    {
        wallPsiDevice(pts, h, rho0, psi);
    }
    else
    {
        wallPsi(pts, h, rho0, psi);
    }
    for (size_t i = 0; i < pts.size(); ++i)
    {
        out[i].position = {pts[i].x, pts[i].y, pts[i].z, pts[i].w};
        out[i].psi = psi[i];
    }
    *count = static_cast<uint32_t>(pts.size());
    return ASPH_OK;
}

}  // namespace asph
