// kernels_step2.hpp -- two time steps per pass on a 3-D box (temporal
// blocking), periodic or with walls, with the intermediate state held in LDS
// at the cell that will pull it.
//
// A block owns a 64 x TY column of output cells and marches through its
// z-range.  Phase A computes the first step on the column grown by one cell
// in x and y, plane by plane; phase B pulls the second step and writes it to
// HBM.  Population q that phase A leaves in grown cell g of plane p is pulled
// by phase B at cell g + c_q (x, y) of plane p + c_q,z, so only the values
// whose destination lies in the tile are kept, and they are kept at that
// destination: slot [q][cell of the tile].  Every plane of every population
// then needs (64+2)*TY words instead of 66*(TY+2), which makes room for a
// taller tile: its grown column re-reads fewer rows, and the re-read rows
// are what the pass is bound by (L1-to-L2 transactions, not arithmetic).
// Phase B reads the slot of its own cell for every q.
//
// At a wall the second step's ghost value of population q at cell x is a
// function of x's own intermediate populations (src/LBM/BoundaryConditions.H:
// 124-189 and src/LBM/BCTypes/), which phase A holds in registers when it
// computes x; it writes the ghost value into the slot that phase B of x reads
// for q, the slot no grown cell pushes into because the pull source lies
// outside the domain.  Per cell the arithmetic is that of k_step (pull,
// collide, wall values), so a pass equals two single steps bit for bit.
//
// Schedule (pair_schedule).  Iteration j computes phase A of plane j+1 (from
// loads issued in the iteration before), issues the loads of plane j+2, and
// after a barrier computes phase B of plane j; a second barrier closes the
// iteration.
#pragma once

#include "kernels_common.hpp"

namespace mfa {

// fn(q) for every population, those with c_z = -1 first, then c_z = 0, then
// c_z = +1 (each class in population order).
template <class L, class Fn>
__device__ __forceinline__ void for_cz(Fn&& fn)
{
    auto cls = [&](auto czc) {
        for_q<L>([&](auto qc) {
            if constexpr (L::c(decltype(qc)::value, 2) == decltype(czc)::value) fn(qc);
        });
    };
    cls(std::integral_constant<int, -1>{});
    cls(std::integral_constant<int, 0>{});
    cls(std::integral_constant<int, 1>{});
}

// Rings of planes by c_z class k = c_z + 1.  Phase B of plane z reads class 0
// from source plane z+1 (written in the same iteration, before the barrier),
// class 1 from plane z (one iteration earlier) and class 2 from plane z-1
// (two iterations earlier): rings of 1, 2 and 3 planes.  The ghost values of
// a wall are written by the cell that pulls them, at its own plane z, into
// the slot of source plane z - c_z; for class 0 that is plane z+1, one
// iteration before phase B of z reads it, while the slot of plane z is still
// being read: with walls class 0 gets a second slot.
// W: the LDS word (double, or the 32-bit word of FP32 storage).
template <class L, int TC, bool WALLS, class W = double>
struct PushRings {
    static_assert(L::D == 3, "three-dimensional lattices only");
    static constexpr int N0 = cz_count<L>(0), N1 = cz_count<L>(1), N2 = cz_count<L>(2);
    static constexpr int S0 = WALLS ? 2 : 1, S1 = 2, S2 = 3;
    static constexpr int PLANES = S0 * N0 + S1 * N1 + S2 * N2;
    W c0[S0][N0][TC];
    W c1[S1][N1][TC];
    W c2[S2][N2][TC];
    // ring slot of one plane in each class (or of the three planes phase B
    // reads); computed once per plane, wave-uniform
    struct Pos {
        int k0, k1, k2;
    };
    // slots of plane p (p >= Z0 - 2; the offset 12 is a multiple of every ring length)
    static __device__ __forceinline__ Pos plane(int p, int Z0)
    {
        const int rel = p - Z0 + 12;
        return Pos{rel % S0, rel % S1, rel % S2};
    }
    // slots phase B of plane z reads: class k from plane z + 1 - k
    static __device__ __forceinline__ Pos pulled_by(int z, int Z0)
    {
        const int rel = z - Z0 + 12;
        return Pos{(rel + 1) % S0, rel % S1, (rel - 1) % S2};
    }
    template <int q>
    __device__ __forceinline__ W& at(const Pos& ps, int cell)
    {
        constexpr int k = L::c(q, 2) + 1, r = cz_rank<L>(q);
        if constexpr (k == 0) return c0[ps.k0][r][cell];
        else if constexpr (k == 1) return c1[ps.k1][r][cell];
        else return c2[ps.k2][r][cell];
    }
};

template <int TY_>
struct PushTile {
    static constexpr int TX = 64, TY = TY_, TC = TX * TY;
    static constexpr int EX = TX + 2, EY = TY + 2, EC = EX * EY;
    static constexpr int THREADS = (EC + 63) / 64 * 64;
    // slots of one tile row, padded by a cell on either side: a wavefront
    // of a grown row pushes its 64 lanes to distinct words
    static constexpr int SX = TX + 2, SC = SX * TY;
    // one block per CU: its waves over the CU's four SIMDs
    static constexpr int WAVES_PER_EU = (THREADS / 64 + 3) / 4;
};

// Grown-column cell (ex, ey) of thread tid in phase A: the tile's own rows
// first, in phase B's order (thread tid < TC computes output cell
// (tid % 64, tid / 64) in both phases), then the rows below and above the
// tile, then the two side columns.  Every wavefront but the last covers one
// row of 64 cells, so its loads are aligned runs.
template <class T>
__device__ __forceinline__ void grown_cell(int tid, int& ex, int& ey)
{
    if (tid < T::TC) {
        ex = tid % T::TX + 1;
        ey = tid / T::TX + 1;
    } else if (tid < T::TC + 2 * T::TX) {
        const int t = tid - T::TC;
        ex = t % T::TX + 1;
        ey = t < T::TX ? 0 : T::TY + 1;
    } else if (tid < T::EC) {
        const int t = tid - T::TC - 2 * T::TX;
        ex = (t & 1) ? T::TX + 1 : 0;
        ey = t >> 1;
    } else {
        // a lane beyond the grown column: the corner cell, whose rows its
        // wavefront loads anyway (its results go nowhere)
        ex = 0;
        ey = 0;
    }
}

// Wall values of one cell: for every population q whose pull source x - c_q
// lies behind a non-periodic face (edge flags lo/hi of the cell), emit(q, v)
// receives the ghost value the reference's BCManager writes there
// (src/LBM/BoundaryConditions.H:124-189; BCTypes/BounceBack.H:56-63,
// OutflowKernel.H:26-38, PressureOutletKernel.H:29-64), computed from s, the
// populations of the cell itself.  The same expressions as wall_value_at,
// with the density and momentum of s that a pressure outlet needs summed
// once per cell rather than once per population.
template <class L, class G, class Emit>
__device__ __forceinline__ void wall_fill(const double (&s)[L::Q], const G& g, const int lo[3],
                                          const int hi[3], bool any_outlet, Emit&& emit)
{
    double rho = 0.0, j0 = 0.0, j1 = 0.0, j2 = 0.0;
    if (any_outlet) {
        for_q<L>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            const double fm = s[m];
            rho += fm;
            j0 += L::c(m, 0) * fm;
            j1 += L::c(m, 1) * fm;
            if constexpr (L::D == 3) j2 += L::c(m, 2) * fm;
        });
    }
    for_q<L>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        int code[3] = {0, 0, 0};
        bool wall = false;
#pragma unroll
        for (int d = 0; d < L::D; ++d) {
            const int c = L::c(q, d);
            if (c == 1 && lo[d]) { code[d] = -1; wall = true; }
            if (c == -1 && hi[d]) { code[d] = 1; wall = true; }
        }
        if (!wall) return;
        const int idx = (code[0] + 1) + 3 * (code[1] + 1) + 9 * (code[2] + 1) * (L::D == 3);
        const int t = g.rtype[idx];
        constexpr int m = L::opp(q);
        double v;
        if (t == MFA_BC_MOVING_WALL || t == MFA_BC_NO_SLIP || t == MFA_BC_INFLOW) {
            v = s[m] + bb_contrib<L, m>(g.rrho[idx], g.ru[idx][0], g.ru[idx][1], g.ru[idx][2]);
        } else if (t == MFA_BC_OUTFLOW) {
            v = s[q];
        } else if (t == MFA_BC_PRESSURE_OUTLET) {
            const double inv_rho = 1.0 / rho;
            const double cu = cdot<L, q>(j0, j1, j2) * inv_rho;
            double uu;
            if constexpr (L::D == 2) uu = (j0 * j0 + j1 * j1) * inv_rho * inv_rho;
            else uu = (j0 * j0 + j1 * j1 + j2 * j2) * inv_rho * inv_rho;
            const double feq_even =
                L::w(q) * g.rrho[idx] * (1.0 + 0.5 * cu * cu * 3.0 * 3.0 - 0.5 * uu * 3.0);
            v = -s[m] + 2.0 * feq_even;
        } else {
            v = __builtin_nan("");
        }
        emit(qc, v);
    });
}

// The march of a pair kernel over planes Z0-1 .. Z1 (phase A) and Z0 .. Z1-1
// (phase B).  Iteration j computes phase A of plane j+1 (from loads issued
// in the iteration before), issues the loads of plane j+2, and after a
// barrier phase B of plane j; a second barrier closes it.  The iterations
// before Z0 (no phase B) and the first with phase B are peeled off the loop,
// and the rim waves (tid >= TC: rows and columns outside the tile) run the
// same schedule without phase B in a loop of their own: then on every path
// into the loop the phase-B stores lie between a load and its use, so the
// compiler's wait for the load leaves those stores in flight.  loadA(P, p)
// and storeA(P, p) are phase A's loads and its collide-and-push of plane p,
// phaseB(z) phase B of plane z; phaseB must not branch around its stores.
template <int TC, class Regs, class LoadA, class StoreA, class PhaseB>
__device__ __forceinline__ void pair_schedule(int tid, int Z0, int Z1, Regs& P, LoadA&& loadA,
                                              StoreA&& storeA, PhaseB&& phaseB)
{
    static_assert(TC % 64 == 0, "the tile/rim split must fall between wavefronts");
    loadA(P, Z0 - 1);
#pragma unroll 1
    for (int p = Z0 - 1; p <= Z0; ++p) {
        storeA(P, p);
        loadA(P, p + 1);
        __syncthreads();
        __syncthreads();
    }
    // uniform over a wavefront; taken from the first lane, the compiler
    // branches on it instead of running both sides under lane masks one
    // after the other
    if (__builtin_amdgcn_readfirstlane(tid) < TC) {
        storeA(P, Z0 + 1);
        if (Z0 + 2 <= Z1) loadA(P, Z0 + 2);
        __syncthreads();
        phaseB(Z0);
        __syncthreads();
#pragma unroll 1
        for (int j = Z0 + 1; j < Z1 - 1; ++j) {
            storeA(P, j + 1);
            loadA(P, j + 2);
            __syncthreads();
            phaseB(j);
            __syncthreads();
        }
        if (Z1 - 1 > Z0) {
            storeA(P, Z1);
            __syncthreads();
            phaseB(Z1 - 1);
        }
    } else {
        storeA(P, Z0 + 1);
        if (Z0 + 2 <= Z1) loadA(P, Z0 + 2);
        __syncthreads();
        __syncthreads();
#pragma unroll 1
        for (int j = Z0 + 1; j < Z1 - 1; ++j) {
            storeA(P, j + 1);
            loadA(P, j + 2);
            __syncthreads();
            __syncthreads();
        }
        if (Z1 - 1 > Z0) {
            storeA(P, Z1);
            __syncthreads();
        }
    }
}

template <class L, bool FORCED, bool WALLS, class T>
__global__ __launch_bounds__(T::THREADS) __attribute__((amdgpu_waves_per_eu(T::WAVES_PER_EU)))
void k_step2p(const double* __restrict__ fin, double* __restrict__ fout, const Geo g, const View vi,
              const View vo, const double omega, const double F0, const double F1, const double F2,
              const int TZ)
{
    __shared__ PushRings<L, T::SC, WALLS> lds;
    // A value whose destination lies outside the tile goes to a dummy row
    // shared by every population and plane, so the writes need no branch.
    __shared__ double dummy[T::SX];
    // The boundary tables, indexed at run time by the ghost code, are read
    // from LDS: indexed in the kernel argument they would be copied to
    // scratch memory.
    __shared__ signed char s_rtype[WALLS ? 27 : 1];
    __shared__ double s_rrho[WALLS ? 27 : 1];
    __shared__ double s_ru[WALLS ? 27 : 1][3];
    using Rings = PushRings<L, T::SC, WALLS>;
    using Pos = typename Rings::Pos;
    const int tid = threadIdx.x;
    GeoRef gr;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        gr.n[d] = g.n[d];
        gr.per[d] = g.per[d];
    }
    gr.walls = g.walls;
    gr.rtype = s_rtype;
    gr.rrho = s_rrho;
    gr.ru = s_ru;
    bool any_outlet = false;
    if constexpr (WALLS) {
#pragma unroll
        for (int i = 0; i < 27; ++i) any_outlet = any_outlet || g.rtype[i] == MFA_BC_PRESSURE_OUTLET;
        if (tid == 0) {
#pragma unroll
            for (int i = 0; i < 27; ++i) {
                s_rtype[i] = g.rtype[i];
                s_rrho[i] = g.rrho[i];
                s_ru[i][0] = g.ru[i][0];
                s_ru[i][1] = g.ru[i][1];
                s_ru[i][2] = g.ru[i][2];
            }
        }
        __syncthreads();
    }
    const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
    const bool px = !WALLS || g.per[0], py = !WALLS || g.per[1], pz = !WALLS || g.per[2];
    // The hardware deals consecutive blocks round-robin to the 8 XCDs.  Give
    // each XCD one contiguous run of tiles (x fastest) instead, so that tiles
    // whose grown columns overlap run in the same L2 and their shared rims
    // and the cache lines their row ends straddle are fetched from HBM once.
    const unsigned nb = gridDim.x * gridDim.y * gridDim.z;
    const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const unsigned per = nb / kXcds, rem = nb % kXcds, xcd = b % kXcds;
    const unsigned t = xcd * per + (xcd < rem ? xcd : rem) + b / kXcds;
    const int X0 = (int)(t % gridDim.x) * T::TX, Y0 = (int)((t / gridDim.x) % gridDim.y) * T::TY;
    const int Z0 = (int)(t / (gridDim.x * gridDim.y)) * TZ;
    const int Z1 = Z0 + TZ < nz ? Z0 + TZ : nz;

    // phase A: one cell of the grown column (periodic images; a grown cell
    // behind a wall is not needed)
    int ex, ey;
    grown_cell<T>(tid, ex, ey);
    bool inA = tid < T::EC;
    int gx = X0 - 1 + ex, gy = Y0 - 1 + ey;
    if (px) gx = (gx % nx + nx) % nx;
    else inA = inA && gx >= 0 && gx < nx;
    if (py) gy = (gy % ny + ny) % ny;
    else inA = inA && gy >= 0 && gy < ny;
    // The in-plane pull sources of the grown cell are fixed for the whole
    // march: byte offsets (x' + y' sy) * 8 for each (c_x, c_y), with x' and y'
    // wrapped (periodic) or clamped (walls; the clamped value is replaced by
    // the wall value), 32-bit because the host has checked that a plane's
    // offsets fit.  The source planes z - c_z are wave-uniform bases, so a
    // load is one scalar base plus one vector offset.
    unsigned srcA[3][3];
    {
        int xs[3], ys[3];
#pragma unroll
        for (int ci = 0; ci < 3; ++ci) {
            int sx = gx - (ci - 1), sy = gy - (ci - 1);
            if (sx < 0) sx = px ? nx - 1 : 0;
            else if (sx >= nx) sx = px ? 0 : nx - 1;
            if (sy < 0) sy = py ? ny - 1 : 0;
            else if (sy >= ny) sy = py ? 0 : ny - 1;
            xs[ci] = sx;
            ys[ci] = sy;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
                srcA[a][b] = (unsigned)(xs[a] + (long long)ys[b] * vi.sy) * (unsigned)sizeof(double);
    }
    // phase B: one output cell
    const int ox = tid % T::TX, oy = tid / T::TX;
    const bool inB = tid < T::TC && X0 + ox < nx && Y0 + oy < ny;
    const int own = ox + 1 + T::SX * oy;    // slot of the output cell
    const unsigned dstB =
        inB ? (unsigned)((X0 + ox) + (long long)(Y0 + oy) * vo.sy) * (unsigned)sizeof(double) : kDropOffset;

    auto has_plane = [&](int p) { return pz || (p >= 0 && p < nz); };
    // edge flags of a cell of the domain at (x, y, z)
    auto edges = [&](int x, int y, int z, int (&lo)[3], int (&hi)[3]) {
        const int pos[3] = {x, y, z};
        bool edge = false;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            lo[d] = !g.per[d] && pos[d] == 0;
            hi[d] = !g.per[d] && pos[d] == g.n[d] - 1;
            edge = edge || lo[d] || hi[d];
        }
        return edge;
    };
    // Without walls every lane of phase A loads and collides: the source
    // offsets of a lane beyond the grown column are wrapped into the box like
    // any other, and push() sends its values to the dummy row.  A branch
    // around the loads or their use would give the compiler a path on which
    // a load is still outstanding at the next iteration's waits.
    auto loadA = [&](double (&P)[L::Q], int p) {
        if constexpr (WALLS)
            if (!inA || !has_plane(p)) return;
        const int pw = (p % nz + nz) % nz;
        const double* plane[3];
#pragma unroll
        for (int ci = 0; ci < 3; ++ci) {
            int sz = pw - (ci - 1);
            if (sz < 0) sz = pz ? nz - 1 : 0;
            else if (sz >= nz) sz = pz ? 0 : nz - 1;
            plane[ci] = fin + (long long)sz * vi.sz;
        }
        // issued plane by plane (z-1, z, z+1): the requests of one source
        // plane leave together (D3Q19 256^3: 0.6 % faster than population
        // order; k_step2p_dev keeps population order, which measured 1 %
        // faster for FP32 words)
        for_cz<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = L::c(q, 0), cy = L::c(q, 1), cz = L::c(q, 2);
            P[q] = *at_bytes(plane[cz + 1] + q * vi.sq, srcA[cx + 1][cy + 1]);
        });
        if constexpr (WALLS) {
            // pulls from behind a wall take the first step's ghost values,
            // from the populations of this cell
            int lo[3], hi[3];
            if (edges(gx, gy, pw, lo, hi)) {
                double s[L::Q];
                for_q<L>([&](auto mc) {
                    constexpr int m = decltype(mc)::value;
                    s[m] = *at_bytes(plane[1] + m * vi.sq, srcA[1][1]);
                });
                wall_fill<L>(s, gr, lo, hi, any_outlet,
                             [&](auto qc, double v) { P[decltype(qc)::value] = v; });
            }
        }
    };
    auto push = [&](auto qc, const Pos& ps, double out) {
        constexpr int q = decltype(qc)::value;
        const int dx = ex - 1 + L::c(q, 0), dy = ey - 1 + L::c(q, 1);
        double* dst = (inA && dx >= 0 && dx < T::TX && dy >= 0 && dy < T::TY)
                          ? &lds.template at<q>(ps, dx + 1 + T::SX * dy)
                          : &dummy[ex];
        *dst = out;
    };
    auto storeA = [&](const double (&P)[L::Q], int p) {
        const Pos pw = Rings::plane(p, Z0);
        if constexpr (WALLS)
            if (!inA || !has_plane(p)) return;
        if constexpr (!WALLS) {
            collide<L, FORCED>(P, omega, F0, F1, F2, [&](auto qc, double out) { push(qc, pw, out); });
        } else {
            double o[L::Q];
            collide<L, FORCED>(P, omega, F0, F1, F2,
                               [&](auto qc, double out) { o[decltype(qc)::value] = out; });
            for_q<L>([&](auto qc) { push(qc, pw, o[decltype(qc)::value]); });
            // the second step's ghost values at a tile cell on an edge go to
            // the slots phase B of this plane reads
            int lo[3], hi[3];
            if (edges(X0 + ox, Y0 + oy, p, lo, hi) && inB) {
                const Pos pr = Rings::pulled_by(p, Z0);
                wall_fill<L>(o, gr, lo, hi, any_outlet, [&](auto qc, double v) {
                    lds.template at<decltype(qc)::value>(pr, own) = v;
                });
            }
        }
    };
    // Phase B runs on the waves of the tile's own rows (tid < TC) for every
    // lane: a lane whose cell lies outside the box reads its (unused) slots
    // and its stores are dropped, so no path skips the stores.
    auto phaseB = [&](int z) {
        const Pos pr = Rings::pulled_by(z, Z0);
        double f[L::Q];
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            f[q] = lds.template at<q>(pr, own);
        });
        double* const orow = fout + (long long)z * vo.sz;
        collide<L, FORCED>(f, omega, F0, F1, F2, [&](auto qc, double out) {
            constexpr int q = decltype(qc)::value;
            store_nt_or_drop(orow + q * vo.sq, dstB, out);
        });
    };

    double P[L::Q];
    pair_schedule<T::TC>(tid, Z0, Z1, P, loadA, storeA, phaseB);
}

// ------------------------------------------------------------ wall shell --
// Two steps on a walled box in two launches.  A cell two or more cells from
// every wall depends, over two steps, only on cells inside the domain: its
// first-step neighbours are at least one cell from the walls, so none of
// their pulls crosses one.  k_step2p run as if the box were periodic gives
// those cells exactly (its wrap-around reaches only the shell), and
// k_step2_faces then recomputes the shell: the cells within one cell of a
// non-periodic face.
//
// A block takes a TA x TB tile of one face's in-plane dimensions (a, b) and
// the face's two shell cells along its normal d.  Phase A computes the first
// step, with k_step's per-cell arithmetic (pull with the wall values of the
// input, collide), for the tile grown by one cell in a and b and for three
// cells along d (the shell and the next cell inward), into LDS as
// [cell][q].  Phase B pulls the second step of each shell cell from there;
// a pull from behind a wall takes the wall value the reference's BCManager
// writes there, computed from the cell's own first-step populations
// (src/LBM/BoundaryConditions.H:124-189 and src/LBM/BCTypes/).  Cells on
// two faces are computed by both, with the same bits.
template <int TA_, int TB_, int THREADS_ = 256>
struct FaceTile {
    static constexpr int TA = TA_, TB = TB_;
    static constexpr int GA = TA + 2, GB = TB + 2, GC = GA * GB * 3;
    static constexpr int OC = 2 * TA * TB;
    static constexpr int THREADS = THREADS_;
};

// The wall faces of a box: normal dimension, side, tiles per in-plane
// dimension and the first block of each face (block ranges, face by face).
struct FaceList {
    int n;
    int dim[6], hi[6], ta[6], tb[6];
    int first[7];
};

template <class L, bool FORCED, class T>
__global__ __launch_bounds__(T::THREADS) void k_step2_faces(const double* __restrict__ fin,
                                                            double* __restrict__ fout, const Geo g,
                                                            const View vi, const View vo,
                                                            const double omega, const double F0,
                                                            const double F1, const double F2,
                                                            const FaceList fl)
{
    static_assert(L::D == 3, "three-dimensional lattices only");
    __shared__ double s1[T::GC][L::Q];
    __shared__ signed char s_rtype[27];
    __shared__ double s_rrho[27];
    __shared__ double s_ru[27][3];
    const int tid = threadIdx.x;
    if (tid < 27) {
        s_rtype[tid] = g.rtype[tid];
        s_rrho[tid] = g.rrho[tid];
        s_ru[tid][0] = g.ru[tid][0];
        s_ru[tid][1] = g.ru[tid][1];
        s_ru[tid][2] = g.ru[tid][2];
    }
    GeoRef gr;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        gr.n[d] = g.n[d];
        gr.per[d] = g.per[d];
    }
    gr.walls = g.walls;
    gr.rtype = s_rtype;
    gr.rrho = s_rrho;
    gr.ru = s_ru;
    // face and tile of this block (wave-uniform)
    const int blk = blockIdx.x;
    int f = 0;
    while (f + 1 < fl.n && blk >= fl.first[f + 1]) ++f;
    const int d = fl.dim[f], hi = fl.hi[f];
    const int da = d == 0 ? 1 : 0, db = d == 2 ? 1 : 2;     // in-plane dimensions, a < b
    const int t = blk - fl.first[f];
    const int a0 = (t % fl.ta[f]) * T::TA, b0 = (t / fl.ta[f]) * T::TB;
    const int na = g.n[da], nb = g.n[db], nd = g.n[d];
    const int d0 = hi ? nd - 3 : 0;                         // first-step slab [d0, d0 + 3)
    __syncthreads();

    // phase A: first step of the grown slab.  Threads walk the slab with x
    // fastest: along a for the y and z faces, along d for the x faces (the
    // three cells of one row side by side in one cache line).
    for (int c = tid; c < T::GC; c += T::THREADS) {
        int ia, ib, id;
        if (d == 0) {
            id = c % 3;
            ia = (c / 3) % T::GA;
            ib = c / (3 * T::GA);
        } else {
            ia = c % T::GA;
            ib = (c / T::GA) % T::GB;
            id = c / (T::GA * T::GB);
        }
        int ca = a0 - 1 + ia, cb = b0 - 1 + ib;
        if (g.per[da]) ca = (ca % na + na) % na;
        else if (ca < 0 || ca >= na) continue;
        if (g.per[db]) cb = (cb % nb + nb) % nb;
        else if (cb < 0 || cb >= nb) continue;
        int pos[3];
        pos[d] = d0 + id;
        pos[da] = ca;
        pos[db] = cb;
        double P[L::Q];
        pull<L, 0, true, 0>(P, fin, pos[0], pos[1], pos[2], gr, vi);
        const int sl = ia + T::GA * (ib + T::GB * id);
        collide<L, FORCED>(P, omega, F0, F1, F2, [&](auto qc, double out) { s1[sl][decltype(qc)::value] = out; });
    }
    __syncthreads();

    // phase B: second step of the shell cells
    for (int o = tid; o < T::OC; o += T::THREADS) {
        int ia, ib, k;
        if (d == 0) {
            k = o % 2;
            ia = (o / 2) % T::TA;
            ib = o / (2 * T::TA);
        } else {
            ia = o % T::TA;
            ib = (o / T::TA) % T::TB;
            k = o / (T::TA * T::TB);
        }
        const int ca = a0 + ia, cb = b0 + ib;
        if (ca >= na || cb >= nb) continue;
        const int cd = hi ? nd - 1 - k : k;
        int pos[3];
        pos[d] = cd;
        pos[da] = ca;
        pos[db] = cb;
        int lo[3], hf[3];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            lo[e] = !g.per[e] && pos[e] == 0;
            hf[e] = !g.per[e] && pos[e] == g.n[e] - 1;
        }
        // grown-slab coordinates of the cell
        const int ga = ia + 1, gb = ib + 1, gd = cd - d0;
        auto slot = [&](int xa, int xb, int xd) { return xa + T::GA * (xb + T::GB * xd); };
        const int own = slot(ga, gb, gd);
        double fq[L::Q];
        for_q<L>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int c3[3] = {L::c(q, 0), L::c(q, 1), L::c(q, 2)};
            // the pull source x - c_q in the grown slab; a source behind the
            // normal face lies outside it and is replaced by the wall value
            int sd = gd - c3[d];
            sd = sd < 0 ? 0 : (sd > 2 ? 2 : sd);
            fq[q] = s1[slot(ga - c3[da], gb - c3[db], sd)][q];
            fix_wall_at<L, q>(fq[q], [&](int m) { return s1[own][m]; }, gr, lo, hf);
        });
        collide_store<L, FORCED>(fq, fout, pos[0] + pos[1] * vo.sy + (long long)pos[2] * vo.sz, vo, omega, F0,
                                 F1, F2);
    }
}

}  // namespace mfa
