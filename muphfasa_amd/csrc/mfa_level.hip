// mfa_level.hip -- several boxes of one lattice on one device: the ghost
// exchange between them (FillBoundary with periodic images) and the
// physical boundary fill of each box at its place in the domain.
//
// The exchange is planned once on the host as rectangular copy regions
// ("tags"), like the copy metadata AMReX keeps for FillBoundary: for every
// destination box b, every source box a and every periodic image s of the
// domain, the cells of a's valid region shifted by s that fall in b's
// one-cell-grown region.  Valid regions tile the domain, so these regions
// are disjoint and lie in b's ghost layer.  One launch copies every
// component of all tags, in chunks of 256 cells.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "kernels_common.hpp"
#include "../../include/mfa_level.h"

namespace mfa {

int field_check(const mfa_field* f, int ncomp_min, int need_grow, const char* what);
void field_range(const mfa_field& f, int D, long long& lo, long long& hi);
Geo geo_for(const int n[3], const mfa_geom_bc* gb, int lat);

struct BoxView {                  // one box's populations, relative to valid cell (0,0,0)
    double* data;
    long long sy, sz, sq;
};

struct CopyTag {                  // a ghost region of box dst imaged from valid cells of box src
    int dst, src;
    int n[3];                     // extents of the region
    int dlo[3];                   // its first cell in dst-local coordinates (a ghost, -1 or n)
    int slo[3];                   // the same cell in src-local coordinates (a valid cell)
};

constexpr int kLevelSlots = 4;

// Each tag is cut into chunks of 256 cells (x fastest); one block per chunk
// finds its tag by binary search over the chunks' prefix sums (the same
// for all its threads) and copies every component of its cells.
constexpr int kFillChunk = 256;

template <class L>
__global__ __launch_bounds__(kFillChunk) void k_level_fill(const CopyTag* __restrict__ tags,
                                                           const long long* __restrict__ cstart,
                                                           const int ntags,
                                                           const BoxView* __restrict__ views)
{
    const long long blk = blockIdx.x + (long long)gridDim.x * blockIdx.y;
    if (blk >= cstart[ntags]) return;
    int lo = 0, hi = ntags - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cstart[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    const CopyTag t = tags[lo];
    const int i = (int)(blk - cstart[lo]) * kFillChunk + (int)threadIdx.x;
    if (i >= t.n[0] * t.n[1] * t.n[2]) return;
    const int rx = i % t.n[0], ry = (i / t.n[0]) % t.n[1], rz = i / (t.n[0] * t.n[1]);
    const BoxView d = views[t.dst], s = views[t.src];
    const long long od = (t.dlo[0] + rx) + (long long)(t.dlo[1] + ry) * d.sy +
                         (long long)(t.dlo[2] + rz) * d.sz;
    const long long os = (t.slo[0] + rx) + (long long)(t.slo[1] + ry) * s.sy +
                         (long long)(t.slo[2] + rz) * s.sz;
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        d.data[q * d.sq + od] = s.data[q * s.sq + os];
    });
}

// The step of every box through its ghost cells (mfa_pull_collide on each
// box) in one launch.  Blocks of 64 x 4 cells of one z-plane; box b owns
// blocks [bstart[b], bstart[b+1]); a block finds its box by binary search
// (the same for all its threads).
template <class L, bool FORCED>
__global__ __launch_bounds__(256) void k_level_step(const BoxView* __restrict__ vin,
                                                    const BoxView* __restrict__ vout,
                                                    const int* __restrict__ boxn,
                                                    const long long* __restrict__ bstart, const int nboxes,
                                                    const Geo g, const double omega, const double F0,
                                                    const double F1, const double F2)
{
    const long long nblk = bstart[nboxes];
    const long long wg = blockIdx.x + (long long)gridDim.x * blockIdx.y;
    if (wg >= nblk) return;
    // Consecutive workgroups go to the XCDs in turn; give each XCD a
    // contiguous range of tiles instead, so that tiles sharing input rows
    // and planes meet in one L2 (a bijection on [0, nblk)).
    const long long per = nblk / kXcds, rem = nblk % kXcds, k = wg % kXcds;
    const long long blk = k * per + (k < rem ? k : rem) + wg / kXcds;
    int lo = 0, hi = nboxes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bstart[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    const int nx = boxn[3 * lo], ny = boxn[3 * lo + 1];
    const int tx = (nx + 63) / 64, ty = (ny + 3) / 4;
    long long r = blk - bstart[lo];
    const int bx = (int)(r % tx);
    r /= tx;
    const int by = (int)(r % ty);
    const int z = (int)(r / ty);
    const int x = bx * 64 + (int)threadIdx.x, y = by * 4 + (int)threadIdx.y;
    if (x >= nx || y >= ny) return;
    const BoxView a = vin[lo], b = vout[lo];
    const View vi{a.sy, a.sz, a.sq}, vo{b.sy, b.sz, b.sq};
    double f[L::Q];
    pull<L, 1, false, 0>(f, a.data, x, y, z, g, vi);
    collide_store<L, FORCED>(f, b.data, x + y * vo.sy + (long long)z * vo.sz, vo, omega, F0, F1, F2);
}

// The neighbours of a box on a periodic tensor-product grid of boxes, seen
// from the box's own frame: entry (cx+1) + 3 (cy+1) + 9 (cz+1) serves the
// cells x - c of the box that fall outside it in the directions given by the
// sign of each c (entry 13 is the box itself).  org is the neighbour's data
// pointer shifted so that org[q sq + x + y sy + z sz] is the neighbour's
// population q at the box-local cell (x, y, z), periodic images included.
struct DirectEntry {
    const double* org;
    long long sy, sz, sq;
};

// The step of every box reading what it pulls from outside itself straight
// from the owning box's valid cells (no ghost exchange, ghost cells never
// read).  NC cells along z per thread, all their loads issued before the
// first collision, as in the one-box step; NT = 1: non-temporal loads of the
// box's own cells, 2: of its own c_x = 0 rows only.  Blocks of 64 x 4 columns; box b
// owns blocks [bstart[b], bstart[b+1]); the 27 entries of the block's box
// are staged in LDS.  The values pulled are those the exchange would have
// put in the ghosts, and the collision is collide_store's, so the result
// equals mfa_level_fill_boundary + mfa_level_step bit for bit.
template <class L, bool FORCED, int NC, int NT>
__global__ __launch_bounds__(256) void k_level_direct(const DirectEntry* __restrict__ table,
                                                      const BoxView* __restrict__ vout,
                                                      const int* __restrict__ boxn,
                                                      const long long* __restrict__ bstart, const int nboxes,
                                                      const double omega, const double F0, const double F1,
                                                      const double F2)
{
    __shared__ DirectEntry T[27];
    const long long nblk = bstart[nboxes];
    const long long wg = blockIdx.x + (long long)gridDim.x * blockIdx.y;
    if (wg >= nblk) return;                       // uniform over the block
    const long long per = nblk / kXcds, rem = nblk % kXcds, k = wg % kXcds;
    const long long blk = k * per + (k < rem ? k : rem) + wg / kXcds;
    int lo = 0, hi = nboxes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bstart[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    const int tid = threadIdx.x + 64 * threadIdx.y;
    if (tid < 27) T[tid] = table[27 * (long long)lo + tid];
    __syncthreads();
    const int nx = boxn[3 * lo], ny = boxn[3 * lo + 1], nz = boxn[3 * lo + 2];
    const int tx = (nx + 63) / 64, ty = (ny + 3) / 4;
    long long r = blk - bstart[lo];
    const int bx = (int)(r % tx);
    r /= tx;
    const int by = (int)(r % ty);
    const int z0 = (int)(r / ty) * NC;
    const int x = bx * 64 + (int)threadIdx.x, y = by * 4 + (int)threadIdx.y;
    if (x >= nx || y >= ny) return;
    const DirectEntry own = T[13];
    double f[NC][L::Q];
#pragma unroll
    for (int kk = 0; kk < NC; ++kk) {
        const int z = z0 + kk;
        if (NC > 1 && z >= nz) break;
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = L::c(q, 0), cy = L::c(q, 1), cz = L::c(q, 2);
            const int sx = x - cx, sy = y - cy, sz = z - cz;
            const int code = (sx < 0 ? 0 : (sx >= nx ? 2 : 1)) + 3 * (sy < 0 ? 0 : (sy >= ny ? 2 : 1)) +
                             9 * (sz < 0 ? 0 : (sz >= nz ? 2 : 1));
            if (__builtin_expect(code == 13, 1)) {
                f[kk][q] = ld<NT == 1 || (NT == 2 && cx == 0)>(&own.org[q * own.sq + sx + sy * own.sy + sz * own.sz]);
            } else {
                const DirectEntry e = T[code];
                f[kk][q] = e.org[q * e.sq + sx + sy * e.sy + sz * e.sz];
            }
        });
    }
    const BoxView b = vout[lo];
    const View vo{b.sy, b.sz, b.sq};
#pragma unroll
    for (int kk = 0; kk < NC; ++kk) {
        const int z = z0 + kk;
        if (NC > 1 && z >= nz) break;
        collide_store<L, FORCED>(f[kk], b.data, x + y * vo.sy + (long long)z * vo.sz, vo, omega, F0, F1, F2);
    }
}

// Physical boundary fill of one box whose valid cells start at lo in a
// domain of dn cells (g.n holds the box extents).  A ghost cell outside the
// domain in a non-periodic dimension takes, for each population q whose
// source g + c_q lies in the box's grown region and in the domain, the wall
// value of its resolved class (src/LBM/BoundaryConditions.H:124-189).
template <class L>
__global__ __launch_bounds__(256) void k_level_bc(double* __restrict__ f, const Geo g, const View v,
                                                  const int lo0, const int lo1, const int lo2,
                                                  const int dn0, const int dn1, const int dn2)
{
    const int gz = (L::D == 3);
    const int x = blockIdx.x * blockDim.x + threadIdx.x - 1;
    const int y = blockIdx.y * blockDim.y + threadIdx.y - 1;
    const int z = (int)blockIdx.z - gz;
    if (x >= g.n[0] + 1 || y >= g.n[1] + 1) return;
    const int p[3] = {x, y, z}, lo[3] = {lo0, lo1, lo2}, dn[3] = {dn0, dn1, dn2};
    int code[3] = {0, 0, 0};
    bool wall = false;
#pragma unroll
    for (int d = 0; d < L::D; ++d) {
        const int P = lo[d] + p[d];
        if (P < 0) code[d] = -1;
        else if (P >= dn[d]) code[d] = 1;
        if (g.per[d]) code[d] = 0;
        if (code[d]) wall = true;
    }
    if (!wall) return;
    const int idx = (code[0] + 1) + 3 * (code[1] + 1) + (L::D == 3 ? 9 * (code[2] + 1) : 0);
    if (g.rtype[idx] == MFA_BC_PERIODIC) return;
    const long long ghost = x + y * v.sy + z * v.sz;
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int s[3] = {x + L::c(q, 0), y + L::c(q, 1), z + L::c(q, 2)};
#pragma unroll
        for (int d = 0; d < L::D; ++d) {
            if (s[d] < -1 || s[d] > g.n[d]) return;              // outside the grown box
            if (lo[d] + s[d] < 0 || lo[d] + s[d] >= dn[d]) return;  // outside the domain
        }
        const long long src = s[0] + s[1] * v.sy + s[2] * v.sz;
        f[q * v.sq + ghost] = wall_value<L, q>(f, src, v, g, idx);
    });
}

template <class Fn>
static int with_lattice_lv(int lat, Fn&& fn)
{
    switch (lat) {
    case MFA_D2Q9: return fn(Lattice<MFA_D2Q9>{});
    case MFA_D3Q15: return fn(Lattice<MFA_D3Q15>{});
    case MFA_D3Q19: return fn(Lattice<MFA_D3Q19>{});
    case MFA_D3Q27: return fn(Lattice<MFA_D3Q27>{});
    default: return fail(MFA_ERR_INVALID, "unknown lattice id");
    }
}

}  // namespace mfa

using namespace mfa;

struct mfa_level {
    int lattice = 0, D = 3, Q = 0, device = 0;
    int dn[3] = {1, 1, 1}, per[3] = {1, 1, 1};
    std::vector<int> lo, n;                 // nboxes x 3
    int nboxes = 0, ntags = 0;
    long long total = 0;
    CopyTag* d_tags = nullptr;
    long long* d_cstart = nullptr;          // fill chunks per tag, prefix sums (ntags + 1)
    long long fill_blocks = 0;
    int* d_boxn = nullptr;                  // nboxes x 3 extents
    long long* d_bstart = nullptr;          // step blocks per box, prefix sums (nboxes + 1)
    long long step_blocks = 0;
    BoxView* d_views = nullptr;             // kLevelSlots x nboxes
    bool bound[kLevelSlots] = {false, false, false, false};
    bool ghosts[kLevelSlots] = {false, false, false, false};   // every box of the slot has ngrow >= 1
    std::vector<BoxView> views[kLevelSlots];
    // byte ranges of the bound boxes, sorted by start
    std::vector<std::pair<long long, long long>> ranges[kLevelSlots];
    // direct step: the boxes form a tensor-product grid on a domain periodic
    // in every dimension; grid position of each box and the box at each
    // grid position
    bool direct = false;
    int gn[3] = {1, 1, 1};
    std::vector<int> gpos, at;              // nboxes x 3; gn[0] gn[1] gn[2]
    long long* d_dstart = nullptr;          // direct-step blocks per box, prefix sums
    long long direct_blocks = 0;
    int direct_nc = 2, direct_nt = 0;       // cells per thread along z, load kind
    DirectEntry* d_direct = nullptr;        // kLevelSlots x nboxes x 27
};

// blocks of the direct step per box (64 x 4 columns of nc cells), prefix sums
static std::vector<long long> direct_starts(const mfa_level* lv, int nc)
{
    std::vector<long long> d(lv->nboxes + 1, 0);
    for (int b = 0; b < lv->nboxes; ++b)
        d[b + 1] = d[b] + (long long)((lv->n[3 * b] + 63) / 64) * ((lv->n[3 * b + 1] + 3) / 4) *
                              ((lv->n[3 * b + 2] + nc - 1) / nc);
    return d;
}

// Do two sets of byte ranges, each sorted by start and disjoint within
// itself, share a byte?
static bool ranges_meet(const std::vector<std::pair<long long, long long>>& a,
                        const std::vector<std::pair<long long, long long>>& b)
{
    std::vector<std::pair<long long, long long>> all(a);
    all.insert(all.end(), b.begin(), b.end());
    std::sort(all.begin(), all.end());
    for (size_t i = 1; i < all.size(); ++i)
        if (all[i].first <= all[i - 1].second) return true;
    return false;
}

static int lattice_q_lv(int lat)
{
    switch (lat) {
    case MFA_D2Q9: return 9;
    case MFA_D3Q15: return 15;
    case MFA_D3Q19: return 19;
    case MFA_D3Q27: return 27;
    default: return -1;
    }
}

static void free_level(mfa_level* lv)
{
    if (lv->d_tags) (void)hipFree(lv->d_tags);
    if (lv->d_cstart) (void)hipFree(lv->d_cstart);
    if (lv->d_boxn) (void)hipFree(lv->d_boxn);
    if (lv->d_bstart) (void)hipFree(lv->d_bstart);
    if (lv->d_views) (void)hipFree(lv->d_views);
    if (lv->d_dstart) (void)hipFree(lv->d_dstart);
    if (lv->d_direct) (void)hipFree(lv->d_direct);
    delete lv;
}

int mfa_level_create(int lattice, const int domain_n[3], const int periodic[3], int nboxes,
                     const int* box_lo, const int* box_n, mfa_level** out)
{
    if (!out) return fail(MFA_ERR_INVALID, "mfa_level_create: null out");
    *out = nullptr;
    const int Q = lattice_q_lv(lattice);
    if (Q < 0) return fail(MFA_ERR_INVALID, "mfa_level_create: unknown lattice");
    if (!domain_n || !periodic || !box_lo || !box_n || nboxes < 1)
        return fail(MFA_ERR_INVALID, "mfa_level_create: null argument or no boxes");
    const int D = lattice == MFA_D2Q9 ? 2 : 3;
    int dn[3], per[3];
    for (int d = 0; d < 3; ++d) {
        dn[d] = d < D ? domain_n[d] : 1;
        per[d] = d < D ? (periodic[d] != 0) : 1;
        if (dn[d] < 1) return fail(MFA_ERR_INVALID, "mfa_level_create: empty domain");
    }
    if (D == 2 && domain_n[2] > 1) return fail(MFA_ERR_INVALID, "mfa_level_create: 2-D lattice needs domain_n[2] == 1");
    long long cells = 0;
    for (int b = 0; b < nboxes; ++b) {
        for (int d = 0; d < 3; ++d) {
            const int l = d < D ? box_lo[3 * b + d] : 0, m = d < D ? box_n[3 * b + d] : 1;
            if (d >= D && (box_lo[3 * b + d] != 0 || box_n[3 * b + d] > 1))
                return fail(MFA_ERR_INVALID, "mfa_level_create: 2-D boxes need lo[2] == 0, n[2] == 1");
            if (m < 1 || l < 0 || l + m > dn[d])
                return fail(MFA_ERR_INVALID, "mfa_level_create: box " + std::to_string(b) + " outside the domain");
        }
        cells += (long long)box_n[3 * b] * box_n[3 * b + 1] * (D == 3 ? box_n[3 * b + 2] : 1);
    }
    for (int a = 0; a < nboxes; ++a)
        for (int b = a + 1; b < nboxes; ++b) {
            bool overlap = true;
            for (int d = 0; d < D; ++d) {
                const int a0 = box_lo[3 * a + d], a1 = a0 + box_n[3 * a + d];
                const int b0 = box_lo[3 * b + d], b1 = b0 + box_n[3 * b + d];
                if (std::max(a0, b0) >= std::min(a1, b1)) overlap = false;
            }
            if (overlap)
                return fail(MFA_ERR_INVALID, "mfa_level_create: boxes " + std::to_string(a) + " and " +
                                                 std::to_string(b) + " overlap");
        }
    if (cells != (long long)dn[0] * dn[1] * dn[2])
        return fail(MFA_ERR_INVALID, "mfa_level_create: the boxes do not cover the domain");

    mfa_level* lv = new mfa_level;
    lv->lattice = lattice;
    lv->D = D;
    lv->Q = Q;
    lv->nboxes = nboxes;
    std::memcpy(lv->dn, dn, sizeof(dn));
    std::memcpy(lv->per, per, sizeof(per));
    lv->lo.resize(3 * nboxes);
    lv->n.resize(3 * nboxes);
    for (int b = 0; b < nboxes; ++b)
        for (int d = 0; d < 3; ++d) {
            lv->lo[3 * b + d] = d < D ? box_lo[3 * b + d] : 0;
            lv->n[3 * b + d] = d < D ? box_n[3 * b + d] : 1;
        }
    // plan the ghost copies
    std::vector<CopyTag> tags;
    const int sr = 1;
    for (int b = 0; b < nboxes; ++b) {
        int G0[3], G1[3];                   // grown region of b, [G0, G1)
        for (int d = 0; d < 3; ++d) {
            G0[d] = lv->lo[3 * b + d] - (d < D ? 1 : 0);
            G1[d] = lv->lo[3 * b + d] + lv->n[3 * b + d] + (d < D ? 1 : 0);
        }
        for (int a = 0; a < nboxes; ++a)
            for (int sz = -sr; sz <= sr; ++sz)
                for (int sy = -sr; sy <= sr; ++sy)
                    for (int sx = -sr; sx <= sr; ++sx) {
                        const int s[3] = {sx, sy, sz};
                        bool ok = true, zero = true;
                        for (int d = 0; d < 3; ++d) {
                            if (s[d] && (d >= D || !per[d])) ok = false;
                            if (s[d]) zero = false;
                        }
                        if (!ok || (zero && a == b)) continue;
                        CopyTag t;
                        t.dst = b;
                        t.src = a;
                        long long count = 1;
                        for (int d = 0; d < 3; ++d) {
                            const int S0 = lv->lo[3 * a + d] + s[d] * dn[d];
                            const int S1 = S0 + lv->n[3 * a + d];
                            const int I0 = std::max(S0, G0[d]), I1 = std::min(S1, G1[d]);
                            if (I0 >= I1) { count = 0; break; }
                            t.n[d] = I1 - I0;
                            t.dlo[d] = I0 - lv->lo[3 * b + d];
                            t.slo[d] = I0 - S0;
                            count *= t.n[d];
                        }
                        if (!count) continue;
                        tags.push_back(t);
                        lv->total += count;
                    }
    }
    lv->ntags = (int)tags.size();
    std::vector<long long> cstart(tags.size() + 1, 0);
    for (size_t t = 0; t < tags.size(); ++t)
        cstart[t + 1] = cstart[t] + ((long long)tags[t].n[0] * tags[t].n[1] * tags[t].n[2] + kFillChunk - 1) /
                                        kFillChunk;
    lv->fill_blocks = cstart.back();
    // blocks of the batched step: 64 x 4 cells of one z-plane
    std::vector<long long> bstart(nboxes + 1, 0);
    for (int b = 0; b < nboxes; ++b)
        bstart[b + 1] = bstart[b] + (long long)((lv->n[3 * b] + 63) / 64) * ((lv->n[3 * b + 1] + 3) / 4) *
                                        lv->n[3 * b + 2];
    lv->step_blocks = bstart[nboxes];
    // a tensor-product grid: the distinct box starts along each dimension
    // cut it, every box spans one interval of each, and every grid position
    // holds exactly one box
    bool all_per = true;
    for (int d = 0; d < D; ++d) all_per = all_per && per[d];
    std::vector<int> cuts[3];
    bool grid = all_per;
    for (int d = 0; d < 3 && grid; ++d) {
        for (int b = 0; b < nboxes; ++b) cuts[d].push_back(lv->lo[3 * b + d]);
        std::sort(cuts[d].begin(), cuts[d].end());
        cuts[d].erase(std::unique(cuts[d].begin(), cuts[d].end()), cuts[d].end());
        lv->gn[d] = (int)cuts[d].size();
        cuts[d].push_back(dn[d]);
    }
    if (grid && (long long)lv->gn[0] * lv->gn[1] * lv->gn[2] != nboxes) grid = false;
    if (grid) {
        lv->gpos.assign(3 * nboxes, 0);
        lv->at.assign(nboxes, -1);
        for (int b = 0; b < nboxes && grid; ++b) {
            int idx = 0, mul = 1;
            for (int d = 0; d < 3; ++d) {
                const int i = (int)(std::lower_bound(cuts[d].begin(), cuts[d].end(), lv->lo[3 * b + d]) -
                                    cuts[d].begin());
                if (cuts[d][i + 1] - cuts[d][i] != lv->n[3 * b + d]) grid = false;
                lv->gpos[3 * b + d] = i;
                idx += i * mul;
                mul *= lv->gn[d];
            }
            if (grid && lv->at[idx] >= 0) grid = false;
            if (grid) lv->at[idx] = b;
        }
    }
    lv->direct = grid;
    const std::vector<long long> dstart = direct_starts(lv, lv->direct_nc);
    lv->direct_blocks = dstart[nboxes];
    hipError_t e = hipGetDevice(&lv->device);
    if (e == hipSuccess && lv->ntags) e = hipMalloc((void**)&lv->d_tags, tags.size() * sizeof(CopyTag));
    if (e == hipSuccess) e = hipMalloc((void**)&lv->d_cstart, cstart.size() * sizeof(long long));
    if (e == hipSuccess) e = hipMalloc((void**)&lv->d_boxn, lv->n.size() * sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&lv->d_bstart, bstart.size() * sizeof(long long));
    if (e == hipSuccess) e = hipMalloc((void**)&lv->d_views, (size_t)kLevelSlots * nboxes * sizeof(BoxView));
    if (e == hipSuccess && lv->ntags)
        e = hipMemcpy(lv->d_tags, tags.data(), tags.size() * sizeof(CopyTag), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(lv->d_cstart, cstart.data(), cstart.size() * sizeof(long long), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(lv->d_boxn, lv->n.data(), lv->n.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(lv->d_bstart, bstart.data(), bstart.size() * sizeof(long long), hipMemcpyHostToDevice);
    if (e == hipSuccess && lv->direct) {
        e = hipMalloc((void**)&lv->d_dstart, dstart.size() * sizeof(long long));
        if (e == hipSuccess)
            e = hipMalloc((void**)&lv->d_direct, (size_t)kLevelSlots * nboxes * 27 * sizeof(DirectEntry));
        if (e == hipSuccess)
            e = hipMemcpy(lv->d_dstart, dstart.data(), dstart.size() * sizeof(long long), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        free_level(lv);
        return fail(e == hipErrorOutOfMemory ? MFA_ERR_NOMEM : MFA_ERR_HIP,
                    std::string("mfa_level_create: ") + hipGetErrorString(e));
    }
    *out = lv;
    return MFA_OK;
}

int mfa_level_destroy(mfa_level* lv)
{
    if (lv) free_level(lv);
    return MFA_OK;
}

long long mfa_level_ghost_cells(const mfa_level* lv) { return lv ? lv->total : 0; }

int mfa_level_bind(mfa_level* lv, int slot, const mfa_field* boxes)
{
    if (!lv || !boxes) return fail(MFA_ERR_INVALID, "mfa_level_bind: null argument");
    if (slot < 0 || slot >= kLevelSlots) return fail(MFA_ERR_INVALID, "mfa_level_bind: slot out of range");
    std::vector<BoxView> v(lv->nboxes);
    bool ghosts = true;
    for (int b = 0; b < lv->nboxes; ++b) {
        const mfa_field* f = &boxes[b];
        int rc = field_check(f, lv->Q, 0, "mfa_level_bind box");
        if (rc) return rc;
        ghosts = ghosts && f->ngrow >= 1;
        for (int d = 0; d < 3; ++d)
            if (f->n[d] != lv->n[3 * b + d])
                return fail(MFA_ERR_INVALID, "mfa_level_bind: box " + std::to_string(b) + " has other extents");
        v[b] = BoxView{(double*)f->data, f->sy, f->sz, f->sq};
    }
    // the boxes of a slot are written together by one launch: they must not
    // share memory
    std::vector<std::pair<long long, long long>> r(lv->nboxes);
    for (int b = 0; b < lv->nboxes; ++b) field_range(boxes[b], lv->D, r[b].first, r[b].second);
    std::sort(r.begin(), r.end());
    for (int b = 1; b < lv->nboxes; ++b)
        if (r[b].first <= r[b - 1].second)
            return fail(MFA_ERR_INVALID, "mfa_level_bind: two boxes of the slot overlap in memory");
    hipError_t e = hipMemcpy(lv->d_views + (size_t)slot * lv->nboxes, v.data(),
                             v.size() * sizeof(BoxView), hipMemcpyHostToDevice);
    if (e == hipSuccess && lv->direct) {
        // the 27 neighbour entries of every box for the direct step
        std::vector<DirectEntry> t((size_t)lv->nboxes * 27);
        for (int b = 0; b < lv->nboxes; ++b)
            for (int code = 0; code < 27; ++code) {
                const int c[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
                int idx = 0, mul = 1;
                long long off[3];
                for (int d = 0; d < 3; ++d) {
                    int i = lv->gpos[3 * b + d] + c[d];
                    long long wrap = 0;
                    if (i < 0) { i += lv->gn[d]; wrap = lv->dn[d]; }
                    if (i >= lv->gn[d]) { i -= lv->gn[d]; wrap = -(long long)lv->dn[d]; }
                    idx += i * mul;
                    mul *= lv->gn[d];
                    off[d] = wrap;                 // lo_b - lo_a is added below
                }
                const int a = lv->at[idx];
                for (int d = 0; d < 3; ++d) off[d] += lv->lo[3 * b + d] - lv->lo[3 * a + d];
                // 2-D boxes have no z neighbours: their cz is always 0
                const BoxView& va = v[a];
                t[(size_t)b * 27 + code] = DirectEntry{va.data + off[0] + off[1] * va.sy + off[2] * va.sz,
                                                       va.sy, va.sz, va.sq};
            }
        e = hipMemcpy(lv->d_direct + (size_t)slot * lv->nboxes * 27, t.data(), t.size() * sizeof(DirectEntry),
                      hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return fail(MFA_ERR_HIP, std::string("mfa_level_bind: ") + hipGetErrorString(e));
    lv->views[slot] = v;
    lv->ranges[slot] = r;
    lv->bound[slot] = true;
    lv->ghosts[slot] = ghosts;
    return MFA_OK;
}

int mfa_level_fill_boundary(const mfa_level* lv, int slot, void* stream)
{
    if (!lv) return fail(MFA_ERR_INVALID, "mfa_level_fill_boundary: null level");
    if (slot < 0 || slot >= kLevelSlots || !lv->bound[slot])
        return fail(MFA_ERR_INVALID, "mfa_level_fill_boundary: slot not bound");
    if (!lv->ghosts[slot])
        return fail(MFA_ERR_INVALID, "mfa_level_fill_boundary: the slot's boxes have no ghost layer "
                                     "(such a slot serves mfa_level_step_direct only)");
    if (!lv->ntags) return MFA_OK;
    return with_lattice_lv(lv->lattice, [&](auto Lc) {
        using L = decltype(Lc);
        const long long nb = lv->fill_blocks;
        const unsigned gx = (unsigned)std::min<long long>(nb, 65535);
        const unsigned gy = (unsigned)((nb + gx - 1) / gx);
        hipLaunchKernelGGL(k_level_fill<L>, dim3(gx, gy), dim3(kFillChunk), 0, (hipStream_t)stream, lv->d_tags,
                           lv->d_cstart, lv->ntags, lv->d_views + (size_t)slot * lv->nboxes);
        MFA_HIP_CHECK(hipGetLastError());
        return (int)MFA_OK;
    });
}

int mfa_level_step(const mfa_level* lv, int slot_in, int slot_out, double omega, const double* force,
                   void* stream)
{
    if (!lv) return fail(MFA_ERR_INVALID, "mfa_level_step: null level");
    for (int s : {slot_in, slot_out})
        if (s < 0 || s >= kLevelSlots || !lv->bound[s])
            return fail(MFA_ERR_INVALID, "mfa_level_step: slot not bound");
    if (slot_in == slot_out) return fail(MFA_ERR_INVALID, "mfa_level_step: slot_in == slot_out");
    if (!lv->ghosts[slot_in])
        return fail(MFA_ERR_INVALID, "mfa_level_step: the input slot's boxes have no ghost layer "
                                     "(such a slot serves mfa_level_step_direct only)");
    if (ranges_meet(lv->ranges[slot_in], lv->ranges[slot_out]))
        return fail(MFA_ERR_INVALID, "mfa_level_step: an input box and an output box overlap in memory");
    return with_lattice_lv(lv->lattice, [&](auto Lc) {
        using L = decltype(Lc);
        Geo g;
        std::memset(&g, 0, sizeof(g));
        const double F0 = force ? force[0] : 0.0, F1 = force ? force[1] : 0.0,
                     F2 = (force && L::D == 3) ? force[2] : 0.0;
        const long long nb = lv->step_blocks;
        const unsigned gx = (unsigned)std::min<long long>(nb, 65535);
        const unsigned gy = (unsigned)((nb + gx - 1) / gx);
        const BoxView* vi = lv->d_views + (size_t)slot_in * lv->nboxes;
        const BoxView* vo = lv->d_views + (size_t)slot_out * lv->nboxes;
        if (force)
            hipLaunchKernelGGL((k_level_step<L, true>), dim3(gx, gy), dim3(64, 4), 0, (hipStream_t)stream, vi, vo,
                               lv->d_boxn, lv->d_bstart, lv->nboxes, g, omega, F0, F1, F2);
        else
            hipLaunchKernelGGL((k_level_step<L, false>), dim3(gx, gy), dim3(64, 4), 0, (hipStream_t)stream, vi,
                               vo, lv->d_boxn, lv->d_bstart, lv->nboxes, g, omega, F0, F1, F2);
        MFA_HIP_CHECK(hipGetLastError());
        return (int)MFA_OK;
    });
}

int mfa_level_direct_available(const mfa_level* lv) { return lv && lv->direct ? 1 : 0; }

int mfa_level_set_direct_launch(mfa_level* lv, int cells_per_thread, int nontemporal)
{
    if (!lv) return fail(MFA_ERR_INVALID, "mfa_level_set_direct_launch: null level");
    if (cells_per_thread != 1 && cells_per_thread != 2 && cells_per_thread != 4)
        return fail(MFA_ERR_INVALID, "mfa_level_set_direct_launch: cells per thread must be 1, 2 or 4");
    if (nontemporal < 0 || nontemporal > 2)
        return fail(MFA_ERR_INVALID, "mfa_level_set_direct_launch: nontemporal must be 0, 1 or 2");
    if (!lv->direct) return fail(MFA_ERR_UNSUPPORTED, "mfa_level_set_direct_launch: no direct step for this level");
    const std::vector<long long> d = direct_starts(lv, cells_per_thread);
    // the block table is read by launches already queued: wait for them
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(lv->d_dstart, d.data(), d.size() * sizeof(long long), hipMemcpyHostToDevice);
    if (e != hipSuccess)
        return fail(MFA_ERR_HIP, std::string("mfa_level_set_direct_launch: ") + hipGetErrorString(e));
    lv->direct_nc = cells_per_thread;
    lv->direct_nt = nontemporal;
    lv->direct_blocks = d[lv->nboxes];
    return MFA_OK;
}

int mfa_level_step_direct(const mfa_level* lv, int slot_in, int slot_out, double omega, const double* force,
                          void* stream)
{
    if (!lv) return fail(MFA_ERR_INVALID, "mfa_level_step_direct: null level");
    if (!lv->direct)
        return fail(MFA_ERR_UNSUPPORTED, "mfa_level_step_direct: needs a domain periodic in every dimension "
                                         "whose boxes form a tensor-product grid (use the exchange and "
                                         "mfa_level_step)");
    for (int s : {slot_in, slot_out})
        if (s < 0 || s >= kLevelSlots || !lv->bound[s])
            return fail(MFA_ERR_INVALID, "mfa_level_step_direct: slot not bound");
    if (slot_in == slot_out) return fail(MFA_ERR_INVALID, "mfa_level_step_direct: slot_in == slot_out");
    if (ranges_meet(lv->ranges[slot_in], lv->ranges[slot_out]))
        return fail(MFA_ERR_INVALID, "mfa_level_step_direct: an input box and an output box overlap in memory");
    return with_lattice_lv(lv->lattice, [&](auto Lc) {
        using L = decltype(Lc);
        const double F0 = force ? force[0] : 0.0, F1 = force ? force[1] : 0.0,
                     F2 = (force && L::D == 3) ? force[2] : 0.0;
        const long long nb = lv->direct_blocks;
        const unsigned gx = (unsigned)std::min<long long>(nb, 65535);
        const unsigned gy = (unsigned)((nb + gx - 1) / gx);
        const DirectEntry* t = lv->d_direct + (size_t)slot_in * lv->nboxes * 27;
        const BoxView* vo = lv->d_views + (size_t)slot_out * lv->nboxes;
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(gx, gy), dim3(64, 4), 0, (hipStream_t)stream, t, vo, lv->d_boxn,
                               lv->d_dstart, lv->nboxes, omega, F0, F1, F2);
        };
        auto with_nt = [&](auto nc) {
            constexpr int NC = decltype(nc)::value;
            const int nt = lv->direct_nt;
            if (force) {
                if (nt == 0) launch(k_level_direct<L, true, NC, 0>);
                else if (nt == 1) launch(k_level_direct<L, true, NC, 1>);
                else launch(k_level_direct<L, true, NC, 2>);
            } else {
                if (nt == 0) launch(k_level_direct<L, false, NC, 0>);
                else if (nt == 1) launch(k_level_direct<L, false, NC, 1>);
                else launch(k_level_direct<L, false, NC, 2>);
            }
        };
        if (lv->direct_nc == 1) with_nt(std::integral_constant<int, 1>{});
        else if (lv->direct_nc == 2) with_nt(std::integral_constant<int, 2>{});
        else with_nt(std::integral_constant<int, 4>{});
        MFA_HIP_CHECK(hipGetLastError());
        return (int)MFA_OK;
    });
}

int mfa_level_apply_bc(const mfa_level* lv, int slot, const mfa_geom_bc* g, void* stream)
{
    if (!lv || !g) return fail(MFA_ERR_INVALID, "mfa_level_apply_bc: null argument");
    if (slot < 0 || slot >= kLevelSlots || !lv->bound[slot])
        return fail(MFA_ERR_INVALID, "mfa_level_apply_bc: slot not bound");
    if (!lv->ghosts[slot])
        return fail(MFA_ERR_INVALID, "mfa_level_apply_bc: the slot's boxes have no ghost layer "
                                     "(such a slot serves mfa_level_step_direct only)");
    if (g->lattice != lv->lattice) return fail(MFA_ERR_INVALID, "mfa_level_apply_bc: lattice differs from the level's");
    for (int d = 0; d < lv->D; ++d)
        if ((g->periodic[d] != 0) != (lv->per[d] != 0))
            return fail(MFA_ERR_INVALID, "mfa_level_apply_bc: periodicity differs from the level's");
    bool walls = false;
    for (int d = 0; d < lv->D; ++d) walls = walls || !lv->per[d];
    if (!walls) return MFA_OK;
    return with_lattice_lv(lv->lattice, [&](auto Lc) {
        using L = decltype(Lc);
        for (int b = 0; b < lv->nboxes; ++b) {
            const int* n = &lv->n[3 * b];
            const int* lo = &lv->lo[3 * b];
            // boxes away from every non-periodic face have no physical ghosts
            bool touches = false;
            for (int d = 0; d < L::D; ++d)
                if (!lv->per[d] && (lo[d] == 0 || lo[d] + n[d] == lv->dn[d])) touches = true;
            if (!touches) continue;
            const Geo geo = geo_for(n, g, lv->lattice);
            const BoxView& bv = lv->views[slot][b];
            const View v{bv.sy, bv.sz, bv.sq};
            const int nx = n[0] + 2;
            const int bx = nx <= 64 ? 64 : (nx <= 128 ? 128 : 256);
            const dim3 block(bx, 256 / bx, 1);
            const dim3 grid((nx + bx - 1) / bx, (n[1] + 2 + block.y - 1) / block.y, n[2] + (L::D == 3 ? 2 : 0));
            hipLaunchKernelGGL(k_level_bc<L>, grid, block, 0, (hipStream_t)stream, bv.data, geo, v, lo[0], lo[1],
                               lo[2], lv->dn[0], lv->dn[1], lv->dn[2]);
            MFA_HIP_CHECK(hipGetLastError());
        }
        return (int)MFA_OK;
    });
}
