// ss_tiles.h -- the tiled fast path: particles kept grouped by spatial tile
// (re-sorted every few steps), and the binning of the next step fused into
// the particle update.
//
// Binning is linear: the per-cell moments the flow solve needs are
//   m_q(cell) = sum over particles p in cell of K_q[angle bin of p],
// the same numbers as contracting the count histogram against the kernels
// (average_stress, stress.rs:60-96; DirectorField, director.rs:36-61; density,
// src/bin/simulation/mod.rs:312-316).  Each K_q[a] is stored as a signed
// 64-bit fixed-point integer, so the sums are exact integers and do not depend
// on the order in which particles are added.  A workgroup owns one tile; it
// accumulates the moments of the cells of its tile plus a one-cell halo in
// LDS, stores the tile's own cells and adds the halo cells to a separate
// inflow array with global atomics.  Particles that drifted farther than the
// halo since the last sort add themselves to the inflow array directly.
#pragma once

#include "ss_kernels.h"

namespace ss {

struct TileGeo {
    int t[3];    // tile extent per axis (cells)
    int nt[3];   // tiles per axis
    int h[3];    // halo per axis (0 when the tile spans the axis)
    int w[3];    // window extent per axis = t + 2h
    int ntiles;
    int ws;      // window cells
    int nxcd;    // workgroup -> tile interleave (XCD count)
};

// Tile of a cell index triple.
__device__ inline int tile_of(const TileGeo& tg, int ix, int iy, int iz) {
    return ((ix / tg.t[0]) * tg.nt[1] + iy / tg.t[1]) * tg.nt[2] + iz / tg.t[2];
}

// Workgroups are dispatched round-robin over the XCDs; give each XCD a
// contiguous block of tiles (a slab of the box) so that the field records and
// moment arrays of neighbouring tiles share one L2.
__device__ inline int tile_for_block(const TileGeo& tg, int b) {
    const int per = (tg.ntiles + tg.nxcd - 1) / tg.nxcd;
    const int x = b % tg.nxcd, k = b / tg.nxcd;
    const int t = x * per + k;
    return t;
}

// Window slot of cell coordinate c on one axis for a tile starting at o, or -1.
__device__ inline int window_coord(int c, int o, int t, int h, int n) {
    if (h == 0) return c - o;  // tile spans the axis
    int d = c - o + 1;
    if (d < 0) d += n;
    if (d >= n) d -= n;
    d -= 1;  // minimum image in [-1, n-2]
    return (d <= t) ? d + 1 : -1;
}

template <typename T, int NM>
struct TileArgs {
    State<T> s;
    const uint32_t* toff;   // tile offsets [ntiles + 1]
    uint64_t n;
    Geo<T> g;
    TileGeo tg;
    UpdateParams<T> P;
    const T* rec;
    int rs;
    uint64_t seed;
    const uint64_t* step;   // device-resident step counter
    uint32_t step_off;      // offset added to *step
    const T* noise;
    const long long* kfix;  // [NM][na] fixed-point moment kernels
    long long* main;        // [ns][NM+1] moments of each tile's own cells
    long long* inflow;      // [ns][NM+1] contributions from other tiles
    uint32_t* hist;         // optional [ns][na] counts (kept distribution)
};

// One tile of one system: move its particles (MOVE) and bin them into the
// next step's moments.  MOVE = false bins the current state (initial
// sampling).  `block` is the workgroup's index within the system's launch.
template <typename T, int NM, bool MOVE, bool EXT_NOISE, bool KEEP>
__device__ inline void tile_update_body(const TileArgs<T, NM>& a, int block, unsigned char* smem_raw) {
    constexpr int NC = NM + 1;
    long long* acc = reinterpret_cast<long long*>(smem_raw);  // [ws][NC]
    long long* kf = acc + (size_t)a.tg.ws * NC;               // [NM][na]
    const Geo<T>& g = a.g;
    const TileGeo& tg = a.tg;
    const int tile = tile_for_block(tg, block);
    if (tile >= tg.ntiles) return;
    const int tz = tile % tg.nt[2], ty = (tile / tg.nt[2]) % tg.nt[1], tx = tile / (tg.nt[2] * tg.nt[1]);
    const int ox = tx * tg.t[0], oy = ty * tg.t[1], oz = tz * tg.t[2];

    for (int k = threadIdx.x; k < tg.ws * NC; k += blockDim.x) acc[k] = 0;
    for (int k = threadIdx.x; k < NM * g.na; k += blockDim.x) kf[k] = a.kfix[k];
    __syncthreads();

    const uint32_t p0 = a.toff[tile], p1 = a.toff[tile + 1];
    for (uint32_t i = p0 + threadIdx.x; i < p1; i += blockDim.x) {
        T nx, ny, nz, ophi, otheta;
        if (MOVE) {
            const uint32_t id = a.s.id[i];
            Noise<T> rv;
            draw_noise<T>(a.P, a.seed, *a.step + a.step_off, id, a.n, a.noise, EXT_NOISE, rv);
            move_particle<T>(a.s, i, id, g, a.P, a.rec, a.rs, rv, nx, ny, nz, ophi, otheta);
        } else {
            nx = a.s.x[i];
            ny = a.s.y[i];
            nz = a.s.z[i];
            ophi = a.s.phi[i];
            otheta = a.s.theta[i];
        }
        const int ix = grid_index(nx, g.gx, g.nx), iy = grid_index(ny, g.gy, g.ny),
                  iz = grid_index(nz, g.gz, g.nz);
        const int ang = angle_of(g, ophi, otheta);
        const int wx = window_coord(ix, ox, tg.t[0], tg.h[0], g.nx);
        const int wy = window_coord(iy, oy, tg.t[1], tg.h[1], g.ny);
        const int wz = window_coord(iz, oz, tg.t[2], tg.h[2], g.nz);
        if (wx >= 0 && wy >= 0 && wz >= 0) {
            long long* dst = acc + (size_t)((wx * tg.w[1] + wy) * tg.w[2] + wz) * NC;
#pragma unroll
            for (int q = 0; q < NM; ++q)
                atomicAdd(reinterpret_cast<unsigned long long*>(dst + q), (unsigned long long)kf[q * g.na + ang]);
            atomicAdd(reinterpret_cast<unsigned long long*>(dst + NM), 1ull);
        } else {
            const int cell = (ix * g.ny + iy) * g.nz + iz;
            long long* dst = a.inflow + (size_t)cell * NC;
#pragma unroll
            for (int q = 0; q < NM; ++q)
                atomicAdd(reinterpret_cast<unsigned long long*>(dst + q), (unsigned long long)kf[q * g.na + ang]);
            atomicAdd(reinterpret_cast<unsigned long long*>(dst + NM), 1ull);
        }
        if (KEEP) {
            const int cell = (ix * g.ny + iy) * g.nz + iz;
            atomicAdd(a.hist + (size_t)cell * g.na + ang, 1u);
        }
    }
    __syncthreads();

    // flush: own cells by plain stores (every cell of the grid is owned by
    // exactly one tile), halo cells with atomics when they received particles
    const int ex = min(tg.t[0], g.nx - ox), ey = min(tg.t[1], g.ny - oy), ez = min(tg.t[2], g.nz - oz);
    for (int slot = threadIdx.x; slot < tg.ws; slot += blockDim.x) {
        const int lz = slot % tg.w[2], ly = (slot / tg.w[2]) % tg.w[1], lx = slot / (tg.w[2] * tg.w[1]);
        const int dx = lx - tg.h[0], dy = ly - tg.h[1], dz = lz - tg.h[2];
        int cx = ox + dx, cy = oy + dy, cz = oz + dz;
        cx = cx < 0 ? cx + g.nx : (cx >= g.nx ? cx - g.nx : cx);
        cy = cy < 0 ? cy + g.ny : (cy >= g.ny ? cy - g.ny : cy);
        cz = cz < 0 ? cz + g.nz : (cz >= g.nz ? cz - g.nz : cz);
        const size_t cell = (size_t)(cx * g.ny + cy) * g.nz + cz;
        const long long* src = acc + (size_t)slot * NC;
        const bool own = dx >= 0 && dx < ex && dy >= 0 && dy < ey && dz >= 0 && dz < ez;
        if (own) {
#pragma unroll
            for (int q = 0; q < NC; ++q) a.main[cell * NC + q] = src[q];
        } else if (src[NM] != 0) {
#pragma unroll
            for (int q = 0; q < NC; ++q)
                atomicAdd(reinterpret_cast<unsigned long long*>(a.inflow + cell * NC + q), (unsigned long long)src[q]);
        }
    }
}

template <typename T, int NM, bool MOVE, bool EXT_NOISE, bool KEEP>
__global__ void __launch_bounds__(256) k_tile_update(TileArgs<T, NM> a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    tile_update_body<T, NM, MOVE, EXT_NOISE, KEEP>(a, blockIdx.x, smem_raw);
}

// Many systems in one persistent launch: work item w of the concatenated
// per-system workgroup ranges [start[j], start[j+1]) is workgroup
// w - start[j] of system j.  The grid strides over all work items.
constexpr int MAX_BATCH = 256;

// Dynamic LDS: [lds_tile bytes for the tile body][njobs + 1 ints of start]; no
// static LDS, so the 64-bit accumulators at the region's start stay 8-byte
// aligned.
template <typename T, int NM, bool MOVE, bool EXT_NOISE>
__global__ void __launch_bounds__(256)
    k_tile_update_batched(const TileArgs<T, NM>* __restrict__ jobs, const int* __restrict__ start, int njobs,
                          int lds_tile) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    int* sstart = reinterpret_cast<int*>(smem_raw + lds_tile);
    for (int k = threadIdx.x; k <= njobs; k += blockDim.x) sstart[k] = start[k];
    __syncthreads();
    const int total = sstart[njobs];
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        int lo = 0, hi = njobs;  // sstart[lo] <= w < sstart[lo + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (sstart[mid] <= w)
                lo = mid;
            else
                hi = mid;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);  // uniform: keeps the job in scalar registers
        const TileArgs<T, NM> a = jobs[lo];
        tile_update_body<T, NM, MOVE, EXT_NOISE, false>(a, w - sstart[lo], smem_raw);
        __syncthreads();
    }
}

// Moments of the next step: main + inflow, fixed point -> working precision,
// into the forward-FFT buffer (imaginary part 0); clears inflow.
template <typename T, int NM>
__global__ void __launch_bounds__(256)
    k_tile_convert(const long long* __restrict__ main, long long* __restrict__ inflow, int ns, double scale,
                   double rho_scale, cplx<T>* __restrict__ spec, T* __restrict__ rho) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ns) return;
    constexpr int NC = NM + 1;
#pragma unroll
    for (int q = 0; q < NM; ++q) {
        const long long v = main[(size_t)c * NC + q] + inflow[(size_t)c * NC + q];
        spec[(size_t)q * ns + c] = cplx<T>{(T)((double)v * scale), T(0)};
        inflow[(size_t)c * NC + q] = 0;
    }
    const long long cntv = main[(size_t)c * NC + NM] + inflow[(size_t)c * NC + NM];
    inflow[(size_t)c * NC + NM] = 0;
    if (rho) rho[c] = (T)((double)cntv * rho_scale);
}

// Particles per tile from the moment arrays' count component (sort after a
// step) -- one thread per tile.
template <int NM>
__global__ void __launch_bounds__(256)
    k_tile_counts(const long long* __restrict__ main, const long long* __restrict__ inflow, int nx, int ny,
                  int nz, TileGeo tg, uint32_t* __restrict__ tcnt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tg.ntiles) return;
    constexpr int NC = NM + 1;
    const int tz = t % tg.nt[2], ty = (t / tg.nt[2]) % tg.nt[1], tx = t / (tg.nt[2] * tg.nt[1]);
    const int ox = tx * tg.t[0], oy = ty * tg.t[1], oz = tz * tg.t[2];
    const int ex = min(tg.t[0], nx - ox), ey = min(tg.t[1], ny - oy), ez = min(tg.t[2], nz - oz);
    long long sum = 0;
    for (int i = 0; i < ex; ++i)
        for (int j = 0; j < ey; ++j)
            for (int k = 0; k < ez; ++k) {
                const size_t cell = (size_t)((ox + i) * ny + oy + j) * nz + oz + k;
                sum += main[cell * NC + NM] + inflow[cell * NC + NM];
            }
    tcnt[t] = (uint32_t)sum;
}

// Particles per tile from the particle state (initial sort).
template <typename T>
__global__ void __launch_bounds__(256)
    k_tile_count_state(State<T> s, uint64_t n, Geo<T> g, TileGeo tg, uint32_t* __restrict__ tcnt) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t t = INVALID_KEY;
    if (i < n)
        t = (uint32_t)tile_of(tg, grid_index(s.x[i], g.gx, g.nx), grid_index(s.y[i], g.gy, g.ny),
                              grid_index(s.z[i], g.gz, g.nz));
    Run r = wave_run(t);
    if (i < n && r.is_head) atomicAdd(tcnt + t, (uint32_t)r.len);
}

// Scatter every particle to its tile segment: runs of equal tiles within a
// wavefront reserve contiguous slots with one atomic.  Used for the first
// sort, when the particles arrive in arbitrary order.
template <typename T>
__device__ inline void copy_particle(const State<T>& a, const State<T>& b, uint64_t i, uint32_t dst) {
    b.x[dst] = a.x[i];
    b.y[dst] = a.y[i];
    b.z[dst] = a.z[i];
    b.phi[dst] = a.phi[i];
    b.theta[dst] = a.theta[i];
    if (a.lx) {
        b.lx[dst] = a.lx[i];
        b.ly[dst] = a.ly[i];
        b.lz[dst] = a.lz[i];
    }
    b.id[dst] = a.id ? a.id[i] : (uint32_t)i;
}

template <typename T>
__device__ inline uint32_t tile_of_particle(const State<T>& a, uint64_t i, const Geo<T>& g, const TileGeo& tg) {
    return (uint32_t)tile_of(tg, grid_index(a.x[i], g.gx, g.nx), grid_index(a.y[i], g.gy, g.ny),
                             grid_index(a.z[i], g.gz, g.nz));
}

template <typename T>
__global__ void __launch_bounds__(256)
    k_tile_permute(State<T> a, State<T> b, uint64_t n, Geo<T> g, TileGeo tg, const uint32_t* __restrict__ toff,
                   uint32_t* __restrict__ cursor) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const uint32_t t = valid ? tile_of_particle(a, i, g, tg) : INVALID_KEY;
    Run r = wave_run(t);
    uint32_t base = 0;
    if (valid && r.is_head) base = atomicAdd(cursor + t, (uint32_t)r.len);
    base = __shfl(base, r.head);
    if (!valid) return;
    copy_particle(a, b, i, toff[t] + base + (uint32_t)(__lane_id() - r.head));
}

// Re-sort of particles that are already grouped by tile up to the few that
// crossed a tile face since the last sort.  Each wavefront owns PW_ITER * 64
// consecutive particles and first counts them per tile in a wave-private LDS
// table, so that a tile's cursor receives one atomic per wavefront and tile
// instead of one per 64 particles; a table that fills up falls back to one
// atomic per run.  A second pass scatters the particles.
constexpr int PW_ITER = 16;
constexpr int PW_TAB = 64;

template <typename T>
__global__ void __launch_bounds__(256)
    k_tile_permute_agg(State<T> a, State<T> b, uint64_t n, Geo<T> g, TileGeo tg, const uint32_t* __restrict__ toff,
                       uint32_t* __restrict__ cursor) {
    __shared__ uint32_t tab_t[4][PW_TAB];
    __shared__ uint32_t tab_c[4][PW_TAB];
    const int wave = threadIdx.x >> 6, lane = __lane_id();
    const uint64_t w0 = ((uint64_t)blockIdx.x * 4 + wave) * (PW_ITER * 64);
    uint32_t* tt = tab_t[wave];
    uint32_t* tc = tab_c[wave];
    int nent = 0;  // wave-uniform
    uint32_t rank[PW_ITER];  // rank within the table entry, or absolute slot (bit 31)
#pragma unroll
    for (int it = 0; it < PW_ITER; ++it) {
        const uint64_t i = w0 + (uint64_t)it * 64 + lane;
        const bool valid = i < n;
        const uint32_t t = valid ? tile_of_particle(a, i, g, tg) : INVALID_KEY;
        Run r = wave_run(t);
        unsigned long long heads = __ballot(r.is_head && valid);
        uint32_t my = 0;
        while (heads) {
            const int h = __builtin_ctzll(heads);
            heads &= heads - 1;
            const uint32_t th = __shfl(t, h);
            const uint32_t len = (uint32_t)__shfl(r.len, h);
            const unsigned long long hit = __ballot(lane < nent && tt[lane] == th);
            uint32_t before;
            int e;
            if (hit) {
                e = __builtin_ctzll(hit);
                before = tc[e];
                if (lane == 0) tc[e] = before + len;
            } else if (nent < PW_TAB) {
                e = nent;
                before = 0;
                if (lane == 0) {
                    tt[e] = th;
                    tc[e] = len;
                }
                nent += 1;
            } else {
                e = -1;
                before = 0;
                if (lane == h) before = atomicAdd(cursor + th, len);
                before = __shfl(before, h);
            }
            __builtin_amdgcn_wave_barrier();
            if (r.head == h) my = (e >= 0) ? (before + (uint32_t)(lane - h)) | ((uint32_t)e << 24)
                                          : (toff[th] + before + (uint32_t)(lane - h)) | 0x80000000u;
        }
        rank[it] = my;
    }
    // one reservation per table entry
    uint32_t ebase = 0;
    if (lane < nent) ebase = atomicAdd(cursor + tt[lane], tc[lane]) + toff[tt[lane]];
#pragma unroll
    for (int it = 0; it < PW_ITER; ++it) {
        const uint64_t i = w0 + (uint64_t)it * 64 + lane;
        const uint32_t rk = rank[it];
        const int e = (int)((rk >> 24) & 0x7F);
        const uint32_t eb = __shfl(ebase, e);
        if (i >= n) continue;
        const uint32_t dst = (rk & 0x80000000u) ? (rk & 0x7FFFFFFFu) : eb + (rk & 0xFFFFFFu);
        copy_particle(a, b, i, dst);
    }
}

}  // namespace ss
