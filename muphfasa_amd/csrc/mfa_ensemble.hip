// mfa_ensemble.hip -- batched engine: many independent 2-D lattices of
// uneven sizes advance together, one kernel launch per time step for the
// whole ensemble.
//
// Each lattice l (nx_l x ny_l cells, its own omega) stores its populations
// contiguously at pop_off_l in the row-interleaved order ([y][q][x]: the Q
// rows of one y side by side).  The work is a
// flat list of tiles (lattice, x0, y0) of TX x TY cells, one cell per thread;
// a fixed grid of blocks walks the list with a block stride (persistent
// blocks, no atomics, deterministic).  Retired (converged) lattices are
// removed from the tile list on the host.  Every lattice shares one resolved
// boundary table (e.g. the lid cavity); an optional per-cell solid mask adds
// half-way bounce-back obstacles inside the domain.  Per-cell arithmetic is
// the single-lattice engine's (FP64: the reference's exact expressions; FP32:
// shifted deviations collided in FP32), so an FP64 ensemble member is
// bit-identical to the same lattice run alone through mfa_step.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mfa_ensemble.h"
#include "kernels_lowprec.hpp"

namespace mfa {

struct LatInfo {
    int nx, ny;
    double omega;
    long long pop_off;   // element offset of the lattice's population block
    long long cell_off;  // cell offset (masks, moments)
};

// near: the tile or its one-cell rim holds a solid cell of the mask (set by
// mfa_ensemble_init from the masks it is given; 1 until then).  Tiles with
// near == 0 skip every mask read: their cells' pulls cannot touch a solid.
struct Tile {
    int lat, x0, y0, near;
};

struct BcTable {
    signed char rtype[27];
    double rrho[27];
    double ru[27][3];
};

template <class L>
__device__ __forceinline__ GeoRef geo_for(const LatInfo& li, const BcTable* bc)
{
    GeoRef g;
    g.n[0] = li.nx;
    g.n[1] = li.ny;
    g.n[2] = 1;
    g.per[0] = g.per[1] = 0;
    g.per[2] = 1;
    g.walls = 1;
    g.rtype = bc->rtype;
    g.rrho = bc->rrho;
    g.ru = bc->ru;
    return g;
}

// Population view of one lattice: row-interleaved, the Q rows of one y side
// by side ([y][q][x]), so that a tile's stores of all Q populations land in
// one contiguous run per row (the single-lattice engine streams this order
// 8-14 % faster than component planes for D2Q9).
template <class L>
__device__ __forceinline__ View pop_view(const LatInfo& li)
{
    return View{(long long)L::Q * li.nx, (long long)L::Q * li.nx * li.ny, (long long)li.nx};
}

// Half-way bounce-back on solid obstacle cells: a population pulled from a
// solid neighbour is the cell's own population pointing into it.
template <class L, class V, class Ld>
__device__ __forceinline__ void obstacle_bb(V (&f)[L::Q], const unsigned char* __restrict__ mask, int x,
                                            int y, int nx, int ny, Ld&& load_self)
{
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int sx = x - L::c(q, 0), sy = y - L::c(q, 1);
        if (sx >= 0 && sx < nx && sy >= 0 && sy < ny && mask[sx + sy * nx])
            f[q] = load_self(L::opp(q));
    });
}

template <class L, class T, bool MASKED>
__global__ __launch_bounds__(256) void k_batch_step(const T* __restrict__ fin, T* __restrict__ fout,
                                                    const unsigned char* __restrict__ masks,
                                                    const Tile* __restrict__ tiles, int ntiles,
                                                    const LatInfo* __restrict__ lats,
                                                    const BcTable* __restrict__ bc)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        const LatInfo li = lats[tl.lat];
        const int x = tl.x0 + threadIdx.x, y = tl.y0 + threadIdx.y;
        if (x >= li.nx || y >= li.ny) continue;
        const View v = pop_view<L>(li);
        const T* a = fin + li.pop_off;
        T* b = fout + li.pop_off;
        const long long self = x + (long long)y * v.sy;      // population index
        const long long cell = x + (long long)y * li.nx;     // mask index
        const unsigned char* mask = MASKED ? masks + li.cell_off : nullptr;
        const bool near = MASKED && tl.near;
        const GeoRef g = geo_for<L>(li, bc);
        if constexpr (std::is_same<T, double>::value) {
            if (near && mask[cell]) {
                for_q<L>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    b[q * v.sq + self] = L::w(q);
                });
                continue;
            }
            double f[L::Q];
            pull<L, 0, true, false>(f, a, x, y, 0, g, v);
            if (near)
                obstacle_bb<L>(f, mask, x, y, li.nx, li.ny, [&](int m) { return a[m * v.sq + self]; });
            collide_store<L, false>(f, b, self, v, li.omega, 0.0, 0.0, 0.0);
        } else {
            if (near && mask[cell]) {
                for_q<L>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    b[q * v.sq + self] = from_f<T>(0.0f);
                });
                continue;
            }
            float d[L::Q];
            pull_dev<L, true>(d, a, x, y, 0, g, v);
            if (near)
                obstacle_bb<L>(d, mask, x, y, li.nx, li.ny, [&](int m) { return to_f(a[m * v.sq + self]); });
            collide_store_dev<L, false>(d, b, self, v, (float)li.omega, 0.f, 0.f, 0.f);
        }
    }
}

// Two steps per launch (temporal blocking).  Pair tiles are PX x PY cells;
// phase A computes the first step on the tile grown by one cell (grown
// cells outside the lattice are not needed: the second step replaces their
// pulls by wall values) into LDS, in the storage type (FP64, or the FP32
// deviation exactly as a single step stores it); phase B pulls the second
// step from LDS.  Wall values and obstacle bounce-back of the second step
// read the cell's own intermediate populations from LDS, through the same
// arithmetic as k_batch_step, so a pair equals two k_batch_step launches
// bit for bit.  near: a solid cell lies within two cells of the tile.
// 32 x 16 tiles: 512 configuration-3 cavities at 0.230 ms per step (FP64),
// against 0.239 (32 x 8), 0.251 (16 x 16) and 0.308 (64 x 4, whose rows
// overhang the 96-, 160- and 224-wide lattices).
constexpr int kPairX = 32, kPairY = 16;
constexpr int kPairEX = kPairX + 2, kPairEY = kPairY + 2, kPairEC = kPairEX * kPairEY;
constexpr int kPairThreads = (kPairEC + 63) / 64 * 64;

template <class L, class T, bool MASKED>
__global__ __launch_bounds__(kPairThreads) void k_batch_step2(const T* __restrict__ fin, T* __restrict__ fout,
                                                             const unsigned char* __restrict__ masks,
                                                             const Tile* __restrict__ tiles, int ntiles,
                                                             const LatInfo* __restrict__ lats,
                                                             const BcTable* __restrict__ bc)
{
    constexpr bool F64 = std::is_same<T, double>::value;
    __shared__ T lds[L::Q][kPairEC];
    const int tid = threadIdx.x;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        const LatInfo li = lats[tl.lat];
        const View v = pop_view<L>(li);
        const T* a = fin + li.pop_off;
        T* b = fout + li.pop_off;
        const unsigned char* mask = MASKED ? masks + li.cell_off : nullptr;
        const bool near = MASKED && tl.near;
        const GeoRef g = geo_for<L>(li, bc);
        if (tid < kPairEC) {                                   // phase A
            const int x = tl.x0 - 1 + tid % kPairEX, y = tl.y0 - 1 + tid / kPairEX;
            if (x >= 0 && x < li.nx && y >= 0 && y < li.ny) {
                const long long self = x + (long long)y * v.sy;
                if (near && mask[x + (long long)y * li.nx]) {
                    for_q<L>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        if constexpr (F64) lds[q][tid] = L::w(q);
                        else lds[q][tid] = from_f<T>(0.0f);
                    });
                } else if constexpr (F64) {
                    double f[L::Q];
                    pull<L, 0, true, false>(f, a, x, y, 0, g, v);
                    if (near)
                        obstacle_bb<L>(f, mask, x, y, li.nx, li.ny, [&](int m) { return a[m * v.sq + self]; });
                    collide<L, false>(f, li.omega, 0.0, 0.0, 0.0, [&](auto qc, double out) {
                        constexpr int q = decltype(qc)::value;
                        lds[q][tid] = out;
                    });
                } else {
                    float d[L::Q], o[L::Q];
                    // cached loads: the rims of neighbouring tiles re-read these rows
                    pull_dev<L, true, T, GeoRef, false>(d, a, x, y, 0, g, v);
                    if (near)
                        obstacle_bb<L>(d, mask, x, y, li.nx, li.ny, [&](int m) { return to_f(a[m * v.sq + self]); });
                    collide_dev<L, false>(d, o, (float)li.omega, 0.f, 0.f, 0.f);
                    for_q<L>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        lds[q][tid] = from_f<T>(o[q]);
                    });
                }
            }
        }
        __syncthreads();
        if (tid < kPairX * kPairY) {                           // phase B
            const int ox = tid % kPairX, oy = tid / kPairX;
            const int x = tl.x0 + ox, y = tl.y0 + oy;
            if (x < li.nx && y < li.ny) {
                const long long self = x + (long long)y * v.sy;
                const int c0 = (ox + 1) + kPairEX * (oy + 1);
                if (near && mask[x + (long long)y * li.nx]) {
                    for_q<L>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        if constexpr (F64) b[q * v.sq + self] = L::w(q);
                        else b[q * v.sq + self] = from_f<T>(0.0f);
                    });
                } else if constexpr (F64) {
                    double f[L::Q];
                    for_q<L>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        f[q] = lds[q][c0 - L::c(q, 0) - kPairEX * L::c(q, 1)];
                    });
                    const int lo[3] = {x == 0, y == 0, 0}, hi[3] = {x == li.nx - 1, y == li.ny - 1, 0};
                    if (lo[0] || lo[1] || hi[0] || hi[1]) {
                        for_q<L>([&](auto qc) {
                            constexpr int q = decltype(qc)::value;
                            fix_wall_at<L, q>(f[q], [&](int m) { return lds[m][c0]; }, g, lo, hi);
                        });
                    }
                    if (near)
                        obstacle_bb<L>(f, mask, x, y, li.nx, li.ny, [&](int m) { return lds[m][c0]; });
                    collide_store<L, false>(f, b, self, v, li.omega, 0.0, 0.0, 0.0);
                } else {
                    float d[L::Q];
                    for_q<L>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        d[q] = to_f(lds[q][c0 - L::c(q, 0) - kPairEX * L::c(q, 1)]);
                    });
                    wall_fix_dev_at<L>(d, [&](int m) { return to_f(lds[m][c0]); }, x, y, 0, g);
                    if (near)
                        obstacle_bb<L>(d, mask, x, y, li.nx, li.ny, [&](int m) { return to_f(lds[m][c0]); });
                    collide_store_dev<L, false>(d, b, self, v, (float)li.omega, 0.f, 0.f, 0.f);
                }
            }
        }
        __syncthreads();
    }
}

// rho, u of every cell (FP64), pulled populations like updateMacroscopic.
template <class L, class T, bool MASKED>
__global__ __launch_bounds__(256) void k_batch_moments(const T* __restrict__ fin,
                                                       const unsigned char* __restrict__ masks,
                                                       const Tile* __restrict__ tiles, int ntiles,
                                                       const LatInfo* __restrict__ lats,
                                                       const BcTable* __restrict__ bc,
                                                       double* __restrict__ rho, double* __restrict__ vel,
                                                       long long total_cells)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        const LatInfo li = lats[tl.lat];
        const int x = tl.x0 + threadIdx.x, y = tl.y0 + threadIdx.y;
        if (x >= li.nx || y >= li.ny) continue;
        const View v = pop_view<L>(li);
        const T* a = fin + li.pop_off;
        const long long self = x + (long long)y * v.sy;      // population index
        const long long cell = li.cell_off + x + (long long)y * li.nx;
        const unsigned char* mask = MASKED ? masks + li.cell_off : nullptr;
        const bool near = MASKED && tl.near;
        if (near && mask[cell - li.cell_off]) {
            rho[cell] = 1.0;
            vel[cell] = 0.0;
            vel[total_cells + cell] = 0.0;
            continue;
        }
        const GeoRef g = geo_for<L>(li, bc);
        double f[L::Q];
        if constexpr (std::is_same<T, double>::value) {
            pull<L, 0, true, false>(f, a, x, y, 0, g, v);
            if (near)
                obstacle_bb<L>(f, mask, x, y, li.nx, li.ny, [&](int m) { return a[m * v.sq + self]; });
        } else {
            float d[L::Q];
            pull_dev<L, true>(d, a, x, y, 0, g, v);
            if (near)
                obstacle_bb<L>(d, mask, x, y, li.nx, li.ny, [&](int m) { return to_f(a[m * v.sq + self]); });
            for_q<L>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                f[q] = L::w(q) + (double)d[q];
            });
        }
        const double F[3] = {0.0, 0.0, 0.0};
        double r, u0, u1, u2;
        moments<L, false>(f, F, r, u0, u1, u2);
        rho[cell] = r;
        vel[cell] = u0;
        vel[total_cells + cell] = u1;
    }
}

// Equilibrium initialisation from per-cell rho/u (warm start) or rest.
template <class L, class T, bool MASKED>
__global__ __launch_bounds__(256) void k_batch_init(T* __restrict__ fout, const unsigned char* __restrict__ masks,
                                                    const Tile* __restrict__ tiles, int ntiles,
                                                    const LatInfo* __restrict__ lats,
                                                    const double* __restrict__ rho,
                                                    const double* __restrict__ vel, long long total_cells)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        const LatInfo li = lats[tl.lat];
        const int x = tl.x0 + threadIdx.x, y = tl.y0 + threadIdx.y;
        if (x >= li.nx || y >= li.ny) continue;
        const View v = pop_view<L>(li);
        const long long self = x + (long long)y * v.sy;
        const long long cell = li.cell_off + x + (long long)y * li.nx;
        const bool solid = MASKED && masks[cell];
        const double r = (rho && !solid) ? rho[cell] : 1.0;
        const double u0 = (vel && !solid) ? vel[cell] : 0.0;
        const double u1 = (vel && !solid) ? vel[total_cells + cell] : 0.0;
        const double usq = usq_of<L>(u0, u1, 0.0);
        T* b = fout + li.pop_off;
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const double fe = feq<L, q>(r, u0, u1, 0.0, usq);
            if constexpr (std::is_same<T, double>::value) b[q * v.sq + self] = fe;
            else b[q * v.sq + self] = from_f<T>((float)(fe - L::w(q)));
        });
    }
}

// One neighbour of a cell along an axis, for the strain-rate stencil: a fluid
// cell (its velocity), a half-way wall at distance 1/2 (the wall's velocity),
// or an open face (no value: zero-gradient).
enum NbKind { NB_FLUID = 0, NB_WALL = 1, NB_OPEN = 2 };

// d u_b / d x_a at a cell from its two neighbours along a (values m, p of the
// component b, kinds km, kp) and its own value c: central differences between
// fluid cells; second-order one-sided differences through a half-way wall
// (the quadratic through the wall value at 1/2 and the two cells on the
// other side); first-order one-sided towards an open face.
__device__ __forceinline__ double ddx(int km, double m, double c, int kp, double p)
{
    if (km == NB_FLUID && kp == NB_FLUID) return 0.5 * (p - m);
    if (km == NB_FLUID && kp == NB_WALL) return (4.0 / 3.0) * p - c - m * (1.0 / 3.0);
    if (km == NB_WALL && kp == NB_FLUID) return p * (1.0 / 3.0) + c - (4.0 / 3.0) * m;
    if (km == NB_WALL && kp == NB_WALL) return p - m;
    if (km == NB_FLUID) return kp == NB_OPEN ? c - m : 0.0;
    if (kp == NB_FLUID) return p - c;
    if (km == NB_WALL) return 2.0 * (c - m);
    if (kp == NB_WALL) return 2.0 * (p - c);
    return 0.0;
}

// First-order Chapman-Enskog start: f_q = feq_q(rho, u) - 3 (tau - 1) w_q rho
// Q_q:S, with Q_q = c_q c_q - I/3 and S the strain rate of the given velocity
// field.  This is the post-collision non-equilibrium part that a BGK lattice
// carries at a smooth flow (tau = 1/omega; the stored populations of a pull
// step are post-collision).  Neighbours across a non-periodic face or in a
// solid are half-way walls (NoSlip at rest; MovingWall and Inflow at their
// resolved velocity); Outflow and PressureOutlet faces are open.  Solid cells
// get the rest equilibrium.
template <class L, class T, bool MASKED>
__global__ __launch_bounds__(256) void k_batch_init_neq(T* __restrict__ fout,
                                                        const unsigned char* __restrict__ masks,
                                                        const Tile* __restrict__ tiles, int ntiles,
                                                        const LatInfo* __restrict__ lats,
                                                        const BcTable* __restrict__ bc,
                                                        const double* __restrict__ rho,
                                                        const double* __restrict__ vel,
                                                        long long total_cells)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        const LatInfo li = lats[tl.lat];
        const int x = tl.x0 + threadIdx.x, y = tl.y0 + threadIdx.y;
        if (x >= li.nx || y >= li.ny) continue;
        const View v = pop_view<L>(li);
        const long long self = x + (long long)y * v.sy;
        const long long cell = li.cell_off + x + (long long)y * li.nx;
        T* b = fout + li.pop_off;
        if (MASKED && masks[cell]) {
            for_q<L>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const double fe = feq<L, q>(1.0, 0.0, 0.0, 0.0, 0.0);
                if constexpr (std::is_same<T, double>::value) b[q * v.sq + self] = fe;
                else b[q * v.sq + self] = from_f<T>((float)(fe - L::w(q)));
            });
            continue;
        }
        const double r = rho[cell];
        const double u0 = vel[cell], u1 = vel[total_cells + cell];
        // g[a][c] = d u_c / d x_a
        double g[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            int kind[2];
            double val[2][2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                int nx = x + (a == 0 ? 2 * s - 1 : 0), ny = y + (a == 1 ? 2 * s - 1 : 0);
                const int n_a = a == 0 ? li.nx : li.ny;
                const int pos = a == 0 ? nx : ny;
                kind[s] = NB_FLUID;
                val[s][0] = val[s][1] = 0.0;
                if (pos < 0 || pos >= n_a) {
                    // face ghost class: code -1/+1 along a, 0 along the other axis
                    const int idx = a == 0 ? (s ? 2 : 0) + 3 : 1 + 3 * (s ? 2 : 0);
                    const int rt = bc->rtype[idx];
                    if (rt == MFA_BC_PERIODIC) {
                        if (a == 0) nx = (nx + li.nx) % li.nx;
                        else ny = (ny + li.ny) % li.ny;
                    } else if (rt == MFA_BC_OUTFLOW || rt == MFA_BC_PRESSURE_OUTLET) {
                        kind[s] = NB_OPEN;
                        continue;
                    } else {
                        kind[s] = NB_WALL;
                        if (rt != MFA_BC_NO_SLIP) {
                            val[s][0] = bc->ru[idx][0];
                            val[s][1] = bc->ru[idx][1];
                        }
                        continue;
                    }
                }
                const long long nc = li.cell_off + nx + (long long)ny * li.nx;
                if (MASKED && masks[nc]) {
                    kind[s] = NB_WALL;
                    continue;
                }
                val[s][0] = vel[nc];
                val[s][1] = vel[total_cells + nc];
            }
            g[a][0] = ddx(kind[0], val[0][0], u0, kind[1], val[1][0]);
            g[a][1] = ddx(kind[0], val[0][1], u1, kind[1], val[1][1]);
        }
        const double sxx = g[0][0], syy = g[1][1], sxy = 0.5 * (g[0][1] + g[1][0]);
        const double tr3 = (sxx + syy) * (1.0 / 3.0);
        const double pre = -3.0 * (1.0 / li.omega - 1.0) * r;
        const double usq = usq_of<L>(u0, u1, 0.0);
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = L::c(q, 0), cy = L::c(q, 1);
            const double qs = cx * cx * sxx + 2.0 * cx * cy * sxy + cy * cy * syy - tr3;
            const double fe = feq<L, q>(r, u0, u1, 0.0, usq) + pre * L::w(q) * qs;
            if constexpr (std::is_same<T, double>::value) b[q * v.sq + self] = fe;
            else b[q * v.sq + self] = from_f<T>((float)(fe - L::w(q)));
        });
    }
}

// near flag of each tile: does the tile or its rim of `rim` cells hold a
// solid cell?
__global__ __launch_bounds__(256) void k_tile_near(const unsigned char* __restrict__ masks,
                                                   const Tile* __restrict__ tiles, int ntiles,
                                                   const LatInfo* __restrict__ lats, int tx, int ty,
                                                   int rim, int* __restrict__ near)
{
    const int hx = tx + 2 * rim, hy = ty + 2 * rim;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        const LatInfo li = lats[tl.lat];
        int any = 0;
        for (int i = threadIdx.x; i < hx * hy; i += blockDim.x) {
            const int x = tl.x0 - rim + i % hx, y = tl.y0 - rim + i / hx;
            if (x >= 0 && x < li.nx && y >= 0 && y < li.ny &&
                masks[li.cell_off + x + (long long)y * li.nx])
                any = 1;
        }
        any = __syncthreads_or(any);
        if (threadIdx.x == 0) near[t] = any;
        __syncthreads();
    }
}

// fa = (fa + fb) / 2 over the cells of every active tile: the average of two
// consecutive states of a warm start (see mfa_ensemble_init_warm).
template <class L, class T>
__global__ __launch_bounds__(256) void k_batch_average(T* __restrict__ fa, const T* __restrict__ fb,
                                                       const Tile* __restrict__ tiles, int ntiles,
                                                       const LatInfo* __restrict__ lats)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        const LatInfo li = lats[tl.lat];
        const int x = tl.x0 + threadIdx.x, y = tl.y0 + threadIdx.y;
        if (x >= li.nx || y >= li.ny) continue;
        const View v = pop_view<L>(li);
        const long long self = li.pop_off + x + (long long)y * v.sy;
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const long long i = self + q * v.sq;
            if constexpr (std::is_same<T, double>::value) fa[i] = 0.5 * (fa[i] + fb[i]);
            else fa[i] = from_f<T>(0.5f * (to_f(fa[i]) + to_f(fb[i])));
        });
    }
}

// Per-tile partial sums of |u - u_prev|^2 and |u|^2 over fluid cells.
template <bool MASKED>
__global__ __launch_bounds__(256) void k_batch_residual(const double* __restrict__ vel,
                                                        const double* __restrict__ vel_prev,
                                                        const unsigned char* __restrict__ masks,
                                                        const Tile* __restrict__ tiles, int ntiles,
                                                        const LatInfo* __restrict__ lats,
                                                        long long total_cells,
                                                        double* __restrict__ part)
{
    __shared__ double sh[4][2];
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const Tile tl = tiles[t];
        const LatInfo li = lats[tl.lat];
        const int x = tl.x0 + threadIdx.x, y = tl.y0 + threadIdx.y;
        double a = 0.0, b = 0.0;
        if (x < li.nx && y < li.ny) {
            const long long cell = li.cell_off + x + (long long)y * li.nx;
            if (!(MASKED && tl.near && masks[cell])) {
                const double u0 = vel[cell], u1 = vel[total_cells + cell];
                const double d0 = u0 - vel_prev[cell], d1 = u1 - vel_prev[total_cells + cell];
                a = d0 * d0 + d1 * d1;
                b = u0 * u0 + u1 * u1;
            }
        }
        const int tid = threadIdx.x + threadIdx.y * blockDim.x;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_down(a, off, 64);
            b += __shfl_down(b, off, 64);
        }
        if ((tid & 63) == 0) {
            sh[tid >> 6][0] = a;
            sh[tid >> 6][1] = b;
        }
        __syncthreads();
        if (tid == 0) {
            const int nw = (blockDim.x * blockDim.y + 63) / 64;
            double sa = 0.0, sb = 0.0;
            for (int w = 0; w < nw; ++w) {
                sa += sh[w][0];
                sb += sh[w][1];
            }
            part[2 * t] = sa;
            part[2 * t + 1] = sb;
        }
        __syncthreads();
    }
}

}  // namespace mfa

using namespace mfa;

struct mfa_ensemble {
    int lattice = MFA_D2Q9, dtype = MFA_F64, nlat = 0, tx = 64, ty = 4, masked = 0, grid = 0;
    std::vector<LatInfo> lats;
    std::vector<int> active;
    std::vector<Tile> tiles;
    long long total_pops = 0, total_cells = 0;
    LatInfo* d_lats = nullptr;
    Tile* d_tiles = nullptr;
    BcTable* d_bc = nullptr;
    double* d_part = nullptr;
    int ntiles = 0, tiles_cap = 0;
    // near flag of every tile of every lattice (active or not), in the order
    // rebuild_tiles enumerates them; empty until masks have been seen
    std::vector<int> near;
    // the same for the pair tiles (kPairX x kPairY, two-cell rim)
    std::vector<Tile> ptiles;
    std::vector<int> pnear;
    Tile* d_ptiles = nullptr;
    int nptiles = 0, ptiles_cap = 0;
};

static thread_local std::string g_ens_err;

static int ens_fail(int code, const std::string& m)
{
    g_ens_err = m;
    return code;
}

#define ENS_HIP(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) return ens_fail(MFA_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

static int rebuild_tiles(mfa_ensemble* e)
{
    e->tiles.clear();
    size_t k = 0;
    const int near_default = e->masked ? 1 : 0;
    for (int l = 0; l < e->nlat; ++l)
        for (int y0 = 0; y0 < e->lats[l].ny; y0 += e->ty)
            for (int x0 = 0; x0 < e->lats[l].nx; x0 += e->tx, ++k)
                if (e->active[l])
                    e->tiles.push_back(Tile{l, x0, y0, k < e->near.size() ? e->near[k] : near_default});
    e->ntiles = (int)e->tiles.size();
    if (e->ntiles > e->tiles_cap) {
        if (e->d_tiles) (void)hipFree(e->d_tiles);
        if (e->d_part) (void)hipFree(e->d_part);
        e->tiles_cap = e->ntiles;
        ENS_HIP(hipMalloc((void**)&e->d_tiles, sizeof(Tile) * e->tiles_cap));
        ENS_HIP(hipMalloc((void**)&e->d_part, sizeof(double) * 2 * e->tiles_cap));
    }
    if (e->ntiles)
        ENS_HIP(hipMemcpy(e->d_tiles, e->tiles.data(), sizeof(Tile) * e->ntiles, hipMemcpyHostToDevice));
    e->ptiles.clear();
    k = 0;
    for (int l = 0; l < e->nlat; ++l)
        for (int y0 = 0; y0 < e->lats[l].ny; y0 += kPairY)
            for (int x0 = 0; x0 < e->lats[l].nx; x0 += kPairX, ++k)
                if (e->active[l])
                    e->ptiles.push_back(Tile{l, x0, y0, k < e->pnear.size() ? e->pnear[k] : near_default});
    e->nptiles = (int)e->ptiles.size();
    if (e->nptiles > e->ptiles_cap) {
        if (e->d_ptiles) (void)hipFree(e->d_ptiles);
        e->ptiles_cap = e->nptiles;
        ENS_HIP(hipMalloc((void**)&e->d_ptiles, sizeof(Tile) * e->ptiles_cap));
    }
    if (e->nptiles)
        ENS_HIP(hipMemcpy(e->d_ptiles, e->ptiles.data(), sizeof(Tile) * e->nptiles, hipMemcpyHostToDevice));
    return MFA_OK;
}

// Near flags of every tile of one tiling (tx x ty tiles, rim cells) from the
// masks; synchronises the stream.
static int near_flags(mfa_ensemble* e, const unsigned char* masks, int tx, int ty, int rim,
                      std::vector<int>& out, hipStream_t st)
{
    std::vector<Tile> all;
    for (int l = 0; l < e->nlat; ++l)
        for (int y0 = 0; y0 < e->lats[l].ny; y0 += ty)
            for (int x0 = 0; x0 < e->lats[l].nx; x0 += tx) all.push_back(Tile{l, x0, y0, 1});
    Tile* d_all = nullptr;
    int* d_near = nullptr;
    const int n = (int)all.size();
    ENS_HIP(hipMalloc((void**)&d_all, sizeof(Tile) * n));
    if (hipMalloc((void**)&d_near, sizeof(int) * n) != hipSuccess) {
        (void)hipFree(d_all);
        return ens_fail(MFA_ERR_HIP, "ensemble: device allocation failed");
    }
    std::vector<int> near(n, 1);
    hipError_t err = hipMemcpyAsync(d_all, all.data(), sizeof(Tile) * n, hipMemcpyHostToDevice, st);
    if (err == hipSuccess) {
        hipLaunchKernelGGL(k_tile_near, dim3(n < 4096 ? n : 4096), dim3(256), 0, st, masks, d_all, n,
                           e->d_lats, tx, ty, rim, d_near);
        err = hipGetLastError();
    }
    if (err == hipSuccess) err = hipMemcpyAsync(near.data(), d_near, sizeof(int) * n, hipMemcpyDeviceToHost, st);
    if (err == hipSuccess) err = hipStreamSynchronize(st);
    (void)hipFree(d_all);
    (void)hipFree(d_near);
    if (err != hipSuccess) return ens_fail(MFA_ERR_HIP, std::string("ensemble: near flags: ") + hipGetErrorString(err));
    out.swap(near);
    return MFA_OK;
}

// Near flags of the step tiles (one-cell rim) and of the pair tiles
// (two-cell rim), then the active tile lists again with the flags in them.
static int compute_near(mfa_ensemble* e, const unsigned char* masks, hipStream_t st)
{
    int rc = near_flags(e, masks, e->tx, e->ty, 1, e->near, st);
    if (rc == MFA_OK) rc = near_flags(e, masks, kPairX, kPairY, 2, e->pnear, st);
    if (rc != MFA_OK) return rc;
    return rebuild_tiles(e);
}

static int grid_for(const mfa_ensemble* e)
{
    int g = e->grid > 0 ? e->grid : e->ntiles;
    if (g > e->ntiles) g = e->ntiles;
    return g < 1 ? 1 : g;
}

template <class Fn>
static int with_type(int dtype, Fn&& fn)
{
    if (dtype == MFA_F64) return fn(double{});
    if (dtype == MFA_F32) return fn(float{});
    return ens_fail(MFA_ERR_UNSUPPORTED, "ensemble: storage must be F64 or F32");
}

extern "C" {

const char* mfa_ensemble_error(void) { return g_ens_err.c_str(); }

mfa_ensemble* mfa_ensemble_create(int lattice, int dtype, int nlat, const int* nx, const int* ny,
                                  const double* omega, const mfa_geom_bc* bc, int tile_x, int tile_y,
                                  int masked)
{
    if (lattice != MFA_D2Q9) { ens_fail(MFA_ERR_UNSUPPORTED, "ensemble: D2Q9 only"); return nullptr; }
    if (dtype != MFA_F64 && dtype != MFA_F32) { ens_fail(MFA_ERR_UNSUPPORTED, "ensemble: F64 or F32"); return nullptr; }
    if (nlat < 1 || !nx || !ny || !omega || !bc) { ens_fail(MFA_ERR_INVALID, "ensemble: bad arguments"); return nullptr; }
    if (tile_x < 1 || tile_y < 1 || tile_x * tile_y > 256 || (tile_x * tile_y) % 64) {
        ens_fail(MFA_ERR_INVALID, "ensemble: tile must hold a multiple of 64 and at most 256 cells");
        return nullptr;
    }
    for (int d = 0; d < 2; ++d)
        if (bc->periodic[d]) { ens_fail(MFA_ERR_UNSUPPORTED, "ensemble: walled lattices only"); return nullptr; }
    auto* e = new mfa_ensemble();
    e->lattice = lattice;
    e->dtype = dtype;
    e->nlat = nlat;
    e->tx = tile_x;
    e->ty = tile_y;
    e->masked = masked ? 1 : 0;
    long long pop = 0, cell = 0;
    for (int l = 0; l < nlat; ++l) {
        if (nx[l] < 1 || ny[l] < 1) { delete e; ens_fail(MFA_ERR_INVALID, "ensemble: empty lattice"); return nullptr; }
        e->lats.push_back(LatInfo{nx[l], ny[l], omega[l], pop, cell});
        pop += 9LL * nx[l] * ny[l];
        cell += (long long)nx[l] * ny[l];
    }
    e->total_pops = pop;
    e->total_cells = cell;
    e->active.assign(nlat, 1);
    BcTable t;
    for (int i = 0; i < 27; ++i) {
        t.rtype[i] = (signed char)bc->resolved_type[i];
        t.rrho[i] = bc->resolved_rho[i];
        for (int d = 0; d < 3; ++d) t.ru[i][d] = bc->resolved_u[i][d];
    }
    bool ok = hipMalloc((void**)&e->d_lats, sizeof(LatInfo) * nlat) == hipSuccess &&
              hipMalloc((void**)&e->d_bc, sizeof(BcTable)) == hipSuccess &&
              hipMemcpy(e->d_lats, e->lats.data(), sizeof(LatInfo) * nlat, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(e->d_bc, &t, sizeof(BcTable), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok || rebuild_tiles(e) != MFA_OK) {
        ens_fail(MFA_ERR_HIP, "ensemble: device allocation failed");
        mfa_ensemble_destroy(e);
        return nullptr;
    }
    return e;
}

void mfa_ensemble_destroy(mfa_ensemble* e)
{
    if (!e) return;
    if (e->d_lats) (void)hipFree(e->d_lats);
    if (e->d_tiles) (void)hipFree(e->d_tiles);
    if (e->d_bc) (void)hipFree(e->d_bc);
    if (e->d_part) (void)hipFree(e->d_part);
    if (e->d_ptiles) (void)hipFree(e->d_ptiles);
    delete e;
}

int mfa_ensemble_info(const mfa_ensemble* e, long long* total_populations, long long* total_cells,
                      long long* pop_offsets, long long* cell_offsets, int* ntiles)
{
    if (!e) return ens_fail(MFA_ERR_INVALID, "null ensemble");
    if (total_populations) *total_populations = e->total_pops;
    if (total_cells) *total_cells = e->total_cells;
    for (int l = 0; l < e->nlat; ++l) {
        if (pop_offsets) pop_offsets[l] = e->lats[l].pop_off;
        if (cell_offsets) cell_offsets[l] = e->lats[l].cell_off;
    }
    if (ntiles) *ntiles = e->ntiles;
    return MFA_OK;
}

int mfa_ensemble_set_active(mfa_ensemble* e, const int* active)
{
    if (!e || !active) return ens_fail(MFA_ERR_INVALID, "null argument");
    for (int l = 0; l < e->nlat; ++l) e->active[l] = active[l] ? 1 : 0;
    return rebuild_tiles(e);
}

int mfa_ensemble_set_grid(mfa_ensemble* e, int blocks)
{
    if (!e || blocks < 0) return ens_fail(MFA_ERR_INVALID, "bad argument");
    e->grid = blocks;
    return MFA_OK;
}

int mfa_ensemble_init(mfa_ensemble* e, void* f, const unsigned char* masks, const double* rho,
                      const double* vel, void* stream)
{
    if (!e || !f) return ens_fail(MFA_ERR_INVALID, "null argument");
    if (e->masked && !masks) return ens_fail(MFA_ERR_INVALID, "ensemble was created with masks");
    if (e->masked) {
        const int rc = compute_near(e, masks, (hipStream_t)stream);
        if (rc != MFA_OK) return rc;
    }
    if (!e->ntiles) return MFA_OK;
    return with_type(e->dtype, [&](auto tv) {
        using T = decltype(tv);
        using L = Lattice<MFA_D2Q9>;
        const dim3 block(e->tx, e->ty);
        if (e->masked)
            hipLaunchKernelGGL((k_batch_init<L, T, true>), dim3(grid_for(e)), block, 0, (hipStream_t)stream, (T*)f, masks, e->d_tiles, e->ntiles, e->d_lats, rho, vel, e->total_cells);
        else
            hipLaunchKernelGGL((k_batch_init<L, T, false>), dim3(grid_for(e)), block, 0, (hipStream_t)stream, (T*)f, masks, e->d_tiles, e->ntiles, e->d_lats, rho, vel, e->total_cells);
        ENS_HIP(hipGetLastError());
        return (int)MFA_OK;
    });
}

int mfa_ensemble_init_warm(mfa_ensemble* e, void* fa, void* fb, const unsigned char* masks,
                           const double* rho, const double* vel, void* stream)
{
    return mfa_ensemble_init_start(e, fa, fb, masks, rho, vel, MFA_START_AVERAGE, stream);
}

int mfa_ensemble_init_start(mfa_ensemble* e, void* fa, void* fb, const unsigned char* masks,
                            const double* rho, const double* vel, int flags, void* stream)
{
    if (!e || !fa) return ens_fail(MFA_ERR_INVALID, "null argument");
    if (flags & ~(MFA_START_NEQ | MFA_START_AVERAGE)) return ens_fail(MFA_ERR_INVALID, "unknown start flags");
    if ((flags & MFA_START_AVERAGE) && (!fb || fa == fb))
        return ens_fail(MFA_ERR_INVALID, "the two-state average needs a second, distinct buffer");
    if ((flags & MFA_START_NEQ) && (!rho || !vel))
        return ens_fail(MFA_ERR_INVALID, "the non-equilibrium start needs rho and u");
    int rc = MFA_OK;
    if (flags & MFA_START_NEQ) {
        if (e->masked && !masks) return ens_fail(MFA_ERR_INVALID, "ensemble was created with masks");
        if (e->masked) rc = compute_near(e, masks, (hipStream_t)stream);
        if (rc == MFA_OK && e->ntiles)
            rc = with_type(e->dtype, [&](auto tv) {
                using T = decltype(tv);
                using L = Lattice<MFA_D2Q9>;
                const dim3 block(e->tx, e->ty);
                if (e->masked)
                    hipLaunchKernelGGL((k_batch_init_neq<L, T, true>), dim3(grid_for(e)), block, 0, (hipStream_t)stream, (T*)fa, masks, e->d_tiles, e->ntiles, e->d_lats, e->d_bc, rho, vel, e->total_cells);
                else
                    hipLaunchKernelGGL((k_batch_init_neq<L, T, false>), dim3(grid_for(e)), block, 0, (hipStream_t)stream, (T*)fa, masks, e->d_tiles, e->ntiles, e->d_lats, e->d_bc, rho, vel, e->total_cells);
                ENS_HIP(hipGetLastError());
                return (int)MFA_OK;
            });
    } else {
        rc = mfa_ensemble_init(e, fa, masks, rho, vel, stream);
    }
    if (rc != MFA_OK || !(flags & MFA_START_AVERAGE)) return rc;
    rc = mfa_ensemble_step(e, fa, fb, masks, stream);
    if (rc != MFA_OK || !e->ntiles) return rc;
    return with_type(e->dtype, [&](auto tv) {
        using T = decltype(tv);
        using L = Lattice<MFA_D2Q9>;
        const dim3 block(e->tx, e->ty);
        hipLaunchKernelGGL((k_batch_average<L, T>), dim3(grid_for(e)), block, 0, (hipStream_t)stream,
                           (T*)fa, (const T*)fb, e->d_tiles, e->ntiles, e->d_lats);
        ENS_HIP(hipGetLastError());
        return (int)MFA_OK;
    });
}

int mfa_ensemble_step(mfa_ensemble* e, const void* fin, void* fout, const unsigned char* masks,
                      void* stream)
{
    if (!e || !fin || !fout || fin == fout) return ens_fail(MFA_ERR_INVALID, "bad populations");
    if (e->masked && !masks) return ens_fail(MFA_ERR_INVALID, "ensemble was created with masks");
    if (!e->ntiles) return MFA_OK;
    return with_type(e->dtype, [&](auto tv) {
        using T = decltype(tv);
        using L = Lattice<MFA_D2Q9>;
        const dim3 block(e->tx, e->ty);
        if (e->masked)
            hipLaunchKernelGGL((k_batch_step<L, T, true>), dim3(grid_for(e)), block, 0, (hipStream_t)stream, (const T*)fin, (T*)fout, masks, e->d_tiles, e->ntiles, e->d_lats, e->d_bc);
        else
            hipLaunchKernelGGL((k_batch_step<L, T, false>), dim3(grid_for(e)), block, 0, (hipStream_t)stream, (const T*)fin, (T*)fout, masks, e->d_tiles, e->ntiles, e->d_lats, e->d_bc);
        ENS_HIP(hipGetLastError());
        return (int)MFA_OK;
    });
}

int mfa_ensemble_step2(mfa_ensemble* e, const void* fin, void* fout, const unsigned char* masks,
                       void* stream)
{
    if (!e || !fin || !fout || fin == fout) return ens_fail(MFA_ERR_INVALID, "bad populations");
    if (e->masked && !masks) return ens_fail(MFA_ERR_INVALID, "ensemble was created with masks");
    if (!e->nptiles) return MFA_OK;
    return with_type(e->dtype, [&](auto tv) {
        using T = decltype(tv);
        using L = Lattice<MFA_D2Q9>;
        const int grid = e->nptiles < 65535 ? e->nptiles : 65535;
        if (e->masked)
            hipLaunchKernelGGL((k_batch_step2<L, T, true>), dim3(grid), dim3(kPairThreads), 0, (hipStream_t)stream, (const T*)fin, (T*)fout, masks, e->d_ptiles, e->nptiles, e->d_lats, e->d_bc);
        else
            hipLaunchKernelGGL((k_batch_step2<L, T, false>), dim3(grid), dim3(kPairThreads), 0, (hipStream_t)stream, (const T*)fin, (T*)fout, masks, e->d_ptiles, e->nptiles, e->d_lats, e->d_bc);
        ENS_HIP(hipGetLastError());
        return (int)MFA_OK;
    });
}

// The launches of nsteps steps: pairs (fa->fb, fb->fa) for whole multiples
// of four steps, so that the state ends in fa after an even number of steps
// and in fb after an odd one, then single steps for the rest.
static int run_launches(mfa_ensemble* e, void* fa, void* fb, const unsigned char* masks, int nsteps,
                        void* stream)
{
    const int quads = nsteps / 4;
    for (int i = 0; i < quads; ++i) {
        int rc = mfa_ensemble_step2(e, fa, fb, masks, stream);
        if (rc == MFA_OK) rc = mfa_ensemble_step2(e, fb, fa, masks, stream);
        if (rc) return rc;
    }
    for (int s = 0; s < nsteps - 4 * quads; ++s) {
        const int rc = mfa_ensemble_step(e, (s & 1) ? fb : fa, (s & 1) ? fa : fb, masks, stream);
        if (rc) return rc;
    }
    return MFA_OK;
}

int mfa_ensemble_run(mfa_ensemble* e, void* fa, void* fb, const unsigned char* masks, int nsteps,
                     int use_graph, void* stream)
{
    // nsteps steps, mostly two per launch (run_launches); the state ends in
    // fa when nsteps is even.  With use_graph the launches are captured once
    // into a hipGraph on a private stream and replayed.
    if (!e || nsteps < 0) return ens_fail(MFA_ERR_INVALID, "bad arguments");
    if (!use_graph || nsteps == 0) return run_launches(e, fa, fb, masks, nsteps, stream);
    // the capture stream, event, graph and executable are destroyed on every
    // way out, errors included
    struct Handles {
        hipStream_t cap = nullptr;
        hipEvent_t ev = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        ~Handles()
        {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            if (ev) (void)hipEventDestroy(ev);
            if (cap) (void)hipStreamDestroy(cap);
        }
    } h;
    ENS_HIP(hipStreamCreateWithFlags(&h.cap, hipStreamNonBlocking));
    ENS_HIP(hipEventCreateWithFlags(&h.ev, hipEventDisableTiming));
    ENS_HIP(hipEventRecord(h.ev, (hipStream_t)stream));
    ENS_HIP(hipStreamWaitEvent(h.cap, h.ev, 0));
    ENS_HIP(hipStreamBeginCapture(h.cap, hipStreamCaptureModeThreadLocal));
    const int rc = run_launches(e, fa, fb, masks, nsteps, h.cap);
    const hipError_t ce = hipStreamEndCapture(h.cap, &h.graph);
    if (rc != MFA_OK) return rc;
    if (ce != hipSuccess) return ens_fail(MFA_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    ENS_HIP(hipGraphInstantiate(&h.exec, h.graph, nullptr, nullptr, 0));
    ENS_HIP(hipGraphLaunch(h.exec, h.cap));
    ENS_HIP(hipEventRecord(h.ev, h.cap));
    ENS_HIP(hipStreamWaitEvent((hipStream_t)stream, h.ev, 0));
    ENS_HIP(hipStreamSynchronize(h.cap));
    return MFA_OK;
}

int mfa_ensemble_moments(mfa_ensemble* e, const void* f, const unsigned char* masks, double* rho,
                         double* vel, void* stream)
{
    if (!e || !f || !rho || !vel) return ens_fail(MFA_ERR_INVALID, "null argument");
    if (e->masked && !masks) return ens_fail(MFA_ERR_INVALID, "ensemble was created with masks");
    if (!e->ntiles) return MFA_OK;
    return with_type(e->dtype, [&](auto tv) {
        using T = decltype(tv);
        using L = Lattice<MFA_D2Q9>;
        const dim3 block(e->tx, e->ty);
        if (e->masked)
            hipLaunchKernelGGL((k_batch_moments<L, T, true>), dim3(grid_for(e)), block, 0, (hipStream_t)stream, (const T*)f, masks, e->d_tiles, e->ntiles, e->d_lats, e->d_bc, rho, vel, e->total_cells);
        else
            hipLaunchKernelGGL((k_batch_moments<L, T, false>), dim3(grid_for(e)), block, 0, (hipStream_t)stream, (const T*)f, masks, e->d_tiles, e->ntiles, e->d_lats, e->d_bc, rho, vel, e->total_cells);
        ENS_HIP(hipGetLastError());
        return (int)MFA_OK;
    });
}

int mfa_ensemble_residual(mfa_ensemble* e, const double* vel, const double* vel_prev,
                          const unsigned char* masks, double* per_lattice, void* stream)
{
    // per_lattice (host, 2*nlat): sum |u-u_prev|^2, sum |u|^2 for active lattices
    if (!e || !vel || !vel_prev || !per_lattice) return ens_fail(MFA_ERR_INVALID, "null argument");
    if (e->masked && !masks) return ens_fail(MFA_ERR_INVALID, "ensemble was created with masks");
    for (int l = 0; l < 2 * e->nlat; ++l) per_lattice[l] = 0.0;
    if (!e->ntiles) return MFA_OK;
    const dim3 block(e->tx, e->ty);
    if (e->masked)
        hipLaunchKernelGGL(k_batch_residual<true>, dim3(grid_for(e)), block, 0, (hipStream_t)stream, vel, vel_prev, masks, e->d_tiles, e->ntiles, e->d_lats, e->total_cells, e->d_part);
    else
        hipLaunchKernelGGL(k_batch_residual<false>, dim3(grid_for(e)), block, 0, (hipStream_t)stream, vel, vel_prev, masks, e->d_tiles, e->ntiles, e->d_lats, e->total_cells, e->d_part);
    ENS_HIP(hipGetLastError());
    std::vector<double> part(2 * (size_t)e->ntiles);
    ENS_HIP(hipMemcpyAsync(part.data(), e->d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, (hipStream_t)stream));
    ENS_HIP(hipStreamSynchronize((hipStream_t)stream));
    for (int t = 0; t < e->ntiles; ++t) {
        per_lattice[2 * e->tiles[t].lat] += part[2 * t];
        per_lattice[2 * e->tiles[t].lat + 1] += part[2 * t + 1];
    }
    return MFA_OK;
}

}  // extern "C"
