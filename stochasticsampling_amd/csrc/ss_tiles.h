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
    long long* main;        // [NM+1][ns] moments of each tile's own cells (count last,
                            // summed only when the density is wanted: P.use_rho)
    long long* inflow;      // [NM+1][ns] contributions from other tiles
    uint32_t* hist;         // optional [ns][na] counts (kept distribution)
    // Every pass that does not re-sort counts, per tile, where its particles
    // went: dcnt[tile][d] for the 27 neighbour directions d (DIR_HOME: the tile
    // itself) and DIR_FAR (outside the tile's window).  A re-sort fused into
    // the next update (dtoff set) writes every particle to state d in the
    // segment of the tile its cell belongs to: the particles of tile S moving
    // in direction d fill the block starting at boff[S][d] (an LDS counter per
    // direction, as only S's workgroup fills it); far particles fill the far
    // block of their tile (farbase, counter farctr).
    uint32_t* dcnt;         // [ntiles][NDC]
    State<T> d;
    const uint32_t* dtoff;  // [ntiles + 1] or nullptr (update in place)
    const uint32_t* boff;   // [ntiles][NDC]
    const uint32_t* farbase;  // [ntiles]
    uint32_t* farctr;       // [ntiles]
    int stage;              // 1: the window's cell records are staged in LDS (8-value records)
    int zz_trace;           // NM = 6: sigma_zz is not summed (flow_point derives it from the trace)
};

// Window-slot tables: entry c of an axis is the window coordinate of cell c
// times the axis's slot stride, or WT_OUT when the cell lies outside the
// window; the slot of a cell is the sum of its three entries, negative when
// outside.
constexpr int WT_OUT = -(1 << 24);

// Neighbour directions of a tile: d = 9 (dx + 1) + 3 (dy + 1) + (dz + 1) for
// the tile at offset (dx, dy, dz), dx, dy, dz in {-1, 0, 1}; DIR_FAR for a cell
// outside the tile's window.  A window-slot table entry carries, above bit
// 16, its axis's share of d (the digit dx + 1 times 9, 3 or 1), so the sum of
// a cell's three entries is its slot + (d << 16).
constexpr int DIR_HOME = 13;
constexpr int DIR_FAR = 27;
constexpr int NDC = 28;  // counters per tile
constexpr int WT_SLOT_BITS = 16;

// Moment kernels in LDS: one row of KS >= NM values per angle bin (an even
// count, so that a row is read as 16-byte pairs).
template <int NM>
constexpr int kf_stride() {
    return (NM + 1) & ~1;
}

// Dynamic LDS of the tile update: window accumulators [ws][nm+1], moment
// kernels [na][KS], the window-slot tables of the three axes (naxes = nx + ny
// + nz ints) followed by the direction counters and block offsets (2 NDC
// ints), and, when staged, the window's cell records [ws][8] of
// rec_bytes / 8 bytes per value (0: not staged).  Every region starts at a
// multiple of 16 bytes.
inline size_t tile_acc_bytes(int ws, int nm) { return ((size_t)ws * (nm + 1) * 8 + 15) / 16 * 16; }
inline size_t tile_lds_bytes(int ws, int nm, int na, int naxes, size_t rec_bytes = 0) {
    const size_t ks = (size_t)((nm + 1) & ~1);
    return tile_acc_bytes(ws, nm) + ks * na * 8 + ((size_t)(naxes + 2 * NDC) * 4 + 15) / 16 * 16 +
           (size_t)ws * rec_bytes;
}


// The particles in slots [q0, q1) of one tile: move them (MOVE) and bin them
// into the next step's moments in the LDS window of the tile; MOVE = false
// bins the current state (initial sampling).  `owner`: this pass stores the
// tile's own cells (every tile has exactly one owning pass per step, which
// also runs for a tile without particles); a pass that does not own the tile
// adds its own cells to the inflow accumulators instead.  kf must hold the
// fixed-point moment kernels; acc is the window.  REUSE: the workgroup
// goes on to another tile and reuses the LDS, so the pass ends with a
// barrier; a workgroup that exits after its one tile does not wait for the
// flush's atomics to complete.
// Grid cell of window slot `slot` of the tile at (ox, oy, oz), periodic;
// own: the cell belongs to the tile.
__device__ inline size_t window_cell(const TileGeo& tg, int nx, int ny, int nz, int ox, int oy, int oz, int slot,
                                     bool* own) {
    const int lz = slot % tg.w[2], ly = (slot / tg.w[2]) % tg.w[1], lx = slot / (tg.w[2] * tg.w[1]);
    const int dx = lx - tg.h[0], dy = ly - tg.h[1], dz = lz - tg.h[2];
    int cx = ox + dx, cy = oy + dy, cz = oz + dz;
    cx = cx < 0 ? cx + nx : (cx >= nx ? cx - nx : cx);
    cy = cy < 0 ? cy + ny : (cy >= ny ? cy - ny : cy);
    cz = cz < 0 ? cz + nz : (cz >= nz ? cz - nz : cz);
    const int ex = min(tg.t[0], nx - ox), ey = min(tg.t[1], ny - oy), ez = min(tg.t[2], nz - oz);
    *own = dx >= 0 && dx < ex && dy >= 0 && dy < ey && dz >= 0 && dz < ez;
    return (size_t)(cx * ny + cy) * nz + cz;
}

// Pointers into LDS typed with its address space, so that the hot loop's
// LDS accesses stay ds_ instructions: with generic pointers the compiler may
// merge an LDS access and a global one on the other side of a branch into a
// single flat access, which waits on both counters.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef long long i64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) long long lds_i64;
typedef __attribute__((address_space(3))) int lds_i32;
typedef __attribute__((address_space(3))) i64x2 lds_i64x2;
typedef __attribute__((address_space(3))) float lds_f32;
typedef __attribute__((address_space(3))) double lds_f64;
typedef __attribute__((address_space(3))) f32x4 lds_f32x4;
typedef __attribute__((address_space(3))) f32x2 lds_f32x2;
typedef __attribute__((address_space(3))) f64x2 lds_f64x2;

// The 8-value cell record of a staged window slot, from LDS.
__device__ inline void staged_record(const float* w, int slot, float r[REC_FULL]) {
    const lds_f32* b = (const lds_f32*)w + slot * REC_SMALL;
    const f32x4 u = *(const lds_f32x4*)b;
    const f32x2 v = *(const lds_f32x2*)(b + 4);
    r[0] = u.x;
    r[1] = u.y;
    r[2] = u.z;
    r[3] = u.w;
    r[4] = v.x;
    r[5] = v.y;
}
__device__ inline void staged_record(const double* w, int slot, double r[REC_SMALL]) {
    const lds_f64* b = (const lds_f64*)w + slot * REC_SMALL;
    const f64x2 u = *(const lds_f64x2*)b;
    const f64x2 v = *(const lds_f64x2*)(b + 2);
    const f64x2 x = *(const lds_f64x2*)(b + 4);
    r[0] = u.x;
    r[1] = u.y;
    r[2] = v.x;
    r[3] = v.y;
    r[4] = x.x;
    r[5] = x.y;
}

template <typename T, int NM, bool MOVE, int NZ, bool KEEP, bool SCATTER, bool FULL = true, bool REUSE = true>
__device__ inline void tile_pass(const TileArgs<T, NM>& a, int tile, uint32_t q0, uint32_t q1, bool owner,
                                 long long* acc, const long long* kf, int* wtab, T* wrec, uint32_t step_off) {
    constexpr int NC = NM + 1;
    constexpr int KS = kf_stride<NM>();
    const Geo<T>& g = a.g;
    const TileGeo& tg = a.tg;
    // the count component (the density) is summed only when the density
    // enters the step (volume_exclusion != 0): nothing else reads it
    const bool counts = FULL && a.P.use_rho;
    // sigma_zz (component 5 of the symmetric 6) from the trace instead
    const bool skip_zz = NM == 6 && a.zz_trace;
    const int tz = tile % tg.nt[2], ty = (tile / tg.nt[2]) % tg.nt[1], tx = tile / (tg.nt[2] * tg.nt[1]);
    const int ox = tx * tg.t[0], oy = ty * tg.t[1], oz = tz * tg.t[2];

    const uint32_t p0 = q0, p1 = q1;
    constexpr bool scatter = MOVE && SCATTER;
    // Software pipeline: the next particle's state is loaded before the
    // current one is updated, so that each wavefront has two iterations of
    // loads in flight.  The first particle's loads are issued before the
    // window is set up, so that they overlap it.
    uint32_t i = p0 + threadIdx.x;
    Loaded<T> cur{};
    uint32_t cid = 0;
    NoiseRec<T> ncur{};  // NZ_REC: the record is prefetched with the state
    if (MOVE && i < p1) {
        cid = at(a.s.id, i);
        load_particle(a.s, i, cur);
        if constexpr (NZ == NZ_REC) ncur = load_noise_rec(a.noise, i);
    }
    for (int k = threadIdx.x; k < tg.ws * NC; k += blockDim.x) acc[k] = 0;
    const int naxes = g.nx + g.ny + g.nz;
    for (int k = threadIdx.x; k < naxes; k += blockDim.x) {
        // (per-axis values by selects: a per-lane index into the tile
        // geometry would be a memory access per lane)
        const int ax = k < g.nx ? 0 : (k < g.nx + g.ny ? 1 : 2);
        const int c = ax == 0 ? k : (ax == 1 ? k - g.nx : k - g.nx - g.ny);
        const int o = ax == 0 ? ox : (ax == 1 ? oy : oz), n = ax == 0 ? g.nx : (ax == 1 ? g.ny : g.nz);
        const int stride = ax == 0 ? tg.w[1] * tg.w[2] : (ax == 1 ? tg.w[2] : 1);
        const int ta = ax == 0 ? tg.t[0] : (ax == 1 ? tg.t[1] : tg.t[2]);
        const int h = ax == 0 ? tg.h[0] : (ax == 1 ? tg.h[1] : tg.h[2]);
        const int w = window_coord(c, o, ta, h, n);
        // the neighbour digit: below the tile, in it (clipped at the grid's
        // end), above it
        const int ext = min(ta, n - o);
        const int dig = w < h ? 0 : (w - h < ext ? 1 : 2);
        const int dmul = ax == 0 ? 9 : (ax == 1 ? 3 : 1);
        wtab[k] = w >= 0 ? w * stride + ((dig * dmul) << WT_SLOT_BITS) : WT_OUT;
    }
    // per direction: particles counted (passes that do not re-sort) or placed
    // (the re-sort), and the re-sort's block offsets
    lds_i32* lcnt = (lds_i32*)wtab + naxes;
    lds_i32* lboff = lcnt + NDC;
    if (threadIdx.x < NDC) {
        lcnt[threadIdx.x] = 0;
        if (scatter) lboff[threadIdx.x] = (int)a.boff[(size_t)tile * NDC + threadIdx.x];
    }
    const lds_i32* wt = (const lds_i32*)wtab;
    const lds_i32* wty = wt + g.nx;
    const lds_i32* wtz = wty + g.ny;
    // the cell records of the window (8-value records): every particle of the
    // tile that is still inside the window reads its record from LDS
    const bool stage = MOVE && a.stage;
    if (stage) {
        for (int slot = threadIdx.x; slot < tg.ws; slot += blockDim.x) {
            bool own;
            const size_t cell = window_cell(tg, g.nx, g.ny, g.nz, ox, oy, oz, slot, &own);
            const T* src = a.rec + cell * REC_SMALL;
            T* dst = wrec + slot * REC_SMALL;
            if constexpr (sizeof(T) == 4) {
                reinterpret_cast<float4*>(dst)[0] = reinterpret_cast<const float4*>(src)[0];
                reinterpret_cast<float4*>(dst)[1] = reinterpret_cast<const float4*>(src)[1];
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    reinterpret_cast<double2*>(dst)[k] = reinterpret_cast<const double2*>(src)[k];
            }
        }
    }
    __syncthreads();

    const uint64_t step = MOVE ? *a.step + step_off : 0;
    // Cell record of a particle: from the staged window, or (outside the
    // window, having drifted since the last re-sort, or records not staged)
    // from global memory; rp: the record's global address, ts: the tile of
    // its cell (re-sorting update).
    auto fetch = [&](const Loaded<T>& q, T r[REC_FULL], const T*& rp, uint32_t& ts, int& dir) {
        const int cx = bin_x(g, q.x), cy = bin_y(g, q.y), cz = bin_z(g, q.z);
        rp = a.rec + (size_t)((cx * g.ny + cy) * g.nz + cz) * a.rs;
        const int sum = wt[cx] + wty[cy] + wtz[cz];
        if constexpr (scatter) {
            ts = (uint32_t)tile_of(tg, cx, cy, cz);
            dir = sum >= 0 ? sum >> WT_SLOT_BITS : DIR_FAR;
        }
        bool from_lds = false;
        if (stage) {
            from_lds = sum >= 0;
            staged_record(wrec, from_lds ? (sum & ((1 << WT_SLOT_BITS) - 1)) : 0, r);
        }
        if (!from_lds) {
            if constexpr (sizeof(T) == 4) {
                gather_record(rp, a.rs, r);
            } else if (stage) {
#pragma unroll
                for (int k = 0; k < REC_SMALL; ++k) r[k] = rp[k];
            }
        }
    };
    // The record of the current particle was fetched at the end of the
    // previous iteration, before that iteration's LDS atomics, so that no
    // iteration starts by waiting for the atomics of the one before.
    T rc[REC_FULL];
    const T* rpc = a.rec;
    uint32_t tsc = 0;
    int dirc = DIR_HOME;
    if (MOVE && i < p1) fetch(cur, rc, rpc, tsc, dirc);
    for (; i < p1; i += blockDim.x) {
        T nx, ny, nz, ophi, otheta;
        Loaded<T> nxt;
        uint32_t nid = 0;
        NoiseRec<T> nnxt{};
        if (MOVE) {
            // unconditional (clamped to the segment's last slot): a branch
            // here would make the wait for the record a wait for everything
            const uint32_t j = min(i + blockDim.x, p1 - 1);
            nid = at(a.s.id, j);
            load_particle(a.s, j, nxt);
            if constexpr (NZ == NZ_REC) nnxt = load_noise_rec(a.noise, j);
            Noise<T> rv;
            if constexpr (NZ == NZ_REC) {
                noise_from_vector<T>(ncur.v, rv);
            } else {
                // seed and step opaque to the compiler here, so that the
                // Philox key schedule derived from them is recomputed by
                // scalar adds in every iteration instead of being hoisted
                // into 20 scalar registers held across the loop
                // (both are uniform: one system per workgroup)
                uint32_t k0 = __builtin_amdgcn_readfirstlane((uint32_t)a.seed);
                uint32_t k1 = __builtin_amdgcn_readfirstlane((uint32_t)(a.seed >> 32));
                uint32_t s0 = __builtin_amdgcn_readfirstlane((uint32_t)step);
                uint32_t s1 = __builtin_amdgcn_readfirstlane((uint32_t)(step >> 32));
                asm volatile("" : "+s"(k0), "+s"(k1), "+s"(s0), "+s"(s1));
                draw_noise<T, NZ>(a.P, ((uint64_t)k1 << 32) | k0, ((uint64_t)s1 << 32) | s0, cid, a.n, a.noise, rv);
            }
            State<T> out = a.s;
            uint32_t o = i;
            if constexpr (scatter) {
                // particles still in this tile take consecutive slots of its
                // block (one LDS reservation per wavefront), the others one
                // slot each of the block of their direction (LDS counters),
                // far particles of their tile's far block
                const bool h = dirc == DIR_HOME;
                const uint64_t m = __ballot(h);
                const int lane = __lane_id();
                if (m != 0) {
                    const int leader = __ffsll((unsigned long long)m) - 1;
                    uint32_t base = 0;
                    if (lane == leader)
                        base = (uint32_t)__hip_atomic_fetch_add(lcnt + DIR_HOME, (int)__popcll(m), __ATOMIC_RELAXED,
                                                                __HIP_MEMORY_SCOPE_WORKGROUP);
                    base = __shfl(base, leader);
                    if (h) o = (uint32_t)lboff[DIR_HOME] + base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
                }
                if (!h) {
                    if (dirc != DIR_FAR)
                        o = (uint32_t)lboff[dirc] +
                            (uint32_t)__hip_atomic_fetch_add(lcnt + dirc, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    else
                        o = a.farbase[tsc] + atomicAdd(a.farctr + tsc, 1u);
                }
                // the counts make every slot distinct and below n; the clamp
                // only keeps the stores inside the arrays whatever the counts
                o = min(o, (uint32_t)a.n - 1u);
                out = a.d;
            }
            // (the slot's first use is the stores after the arithmetic, so
            // the reservation's latency overlaps it)
            if (sizeof(T) == 4 || stage)
                move_loaded<T, FULL>(out, o, cur, a.P, rc, rv, nx, ny, nz, ophi, otheta);
            else
                move_loaded<T, FULL>(out, o, cur, a.P, rpc, rv, nx, ny, nz, ophi, otheta);
            if constexpr (scatter) at(out.id, o) = cid;
        } else {
            nx = at(a.s.x, i);
            ny = at(a.s.y, i);
            nz = at(a.s.z, i);
            ophi = at(a.s.phi, i);
            otheta = at(a.s.theta, i);
        }
        const int ix = bin_x(g, nx), iy = bin_y(g, ny), iz = bin_z(g, nz);
        const int ang = angle_of(g, ophi, otheta);
        const int wsum = wt[ix] + wty[iy] + wtz[iz];
        const int wslot = wsum >= 0 ? wsum & ((1 << WT_SLOT_BITS) - 1) : -1;
        if constexpr (!scatter) {
            // where the particle went, for the next re-sort
            const int d = wsum >= 0 ? wsum >> WT_SLOT_BITS : DIR_FAR;
            const bool h = d == DIR_HOME;
            const uint64_t m = __ballot(h);
            if (m != 0 && __lane_id() == __ffsll((unsigned long long)m) - 1)
                __hip_atomic_fetch_add(lcnt + DIR_HOME, (int)__popcll(m), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (!h) __hip_atomic_fetch_add(lcnt + d, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        const lds_i64* kr = (const lds_i64*)kf + ang * KS;
        i64x2 kv[KS / 2];
        if (wslot >= 0) {
#pragma unroll
            for (int q = 0; q < KS / 2; ++q) kv[q] = ((const lds_i64x2*)kr)[q];
        }
        if (MOVE) {
            fetch(nxt, rc, rpc, tsc, dirc);
            cur = nxt;
            cid = nid;
            if constexpr (NZ == NZ_REC) ncur = nnxt;
        }
        if (wslot >= 0) {
            lds_i64* dst = (lds_i64*)acc + wslot * NC;
#pragma unroll
            for (int q = 0; q < NM; ++q)
                if (!(skip_zz && q == 5))
                    __hip_atomic_fetch_add(dst + q, (q & 1) ? kv[q / 2].y : kv[q / 2].x, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
            if (counts) __hip_atomic_fetch_add(dst + NM, 1ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            const int cell = (ix * g.ny + iy) * g.nz + iz;
            long long* dst = a.inflow + cell;
#pragma unroll
            for (int q = 0; q < NM; ++q)
                if (!(skip_zz && q == 5))
                    atomicAdd(reinterpret_cast<unsigned long long*>(dst + (size_t)q * g.ns), (unsigned long long)kr[q]);
            if (counts) atomicAdd(reinterpret_cast<unsigned long long*>(dst + (size_t)NM * g.ns), 1ull);
        }
        if (KEEP) {
            const int cell = (ix * g.ny + iy) * g.nz + iz;
            atomicAdd(a.hist + (size_t)cell * g.na + ang, 1u);
        }
    }
    __syncthreads();

    if (!scatter && threadIdx.x < NDC) a.dcnt[(size_t)tile * NDC + threadIdx.x] = (uint32_t)lcnt[threadIdx.x];
    // flush: own cells by plain stores when this pass owns the tile (every
    // cell of the grid is owned by exactly one tile, every tile by exactly
    // one pass), other cells with atomics when they received particles (a
    // cell whose sums are all zero is skipped whether or not it received
    // any: adding zeros changes nothing)
    for (int slot = threadIdx.x; slot < tg.ws; slot += blockDim.x) {
        bool own;
        const size_t cell = window_cell(tg, g.nx, g.ny, g.nz, ox, oy, oz, slot, &own);
        const long long* src = acc + (size_t)slot * NC;
        if (own && owner) {
#pragma unroll
            for (int q = 0; q < NM; ++q)
                if (!(skip_zz && q == 5)) a.main[(size_t)q * g.ns + cell] = src[q];
            if (counts) a.main[(size_t)NM * g.ns + cell] = src[NM];
        } else {
            long long any = counts ? src[NM] : 0;
#pragma unroll
            for (int q = 0; q < NM; ++q) any |= src[q];
            if (any != 0) {
#pragma unroll
                for (int q = 0; q < NM; ++q)
                    if (!(skip_zz && q == 5))
                        atomicAdd(reinterpret_cast<unsigned long long*>(a.inflow + (size_t)q * g.ns + cell),
                                  (unsigned long long)src[q]);
                if (counts)
                    atomicAdd(reinterpret_cast<unsigned long long*>(a.inflow + (size_t)NM * g.ns + cell),
                              (unsigned long long)src[NM]);
            }
        }
    }
    if constexpr (REUSE) __syncthreads();
}

// Moment kernels [NM][na] into LDS rows [na][KS] (visible after the first
// barrier of tile_pass).
template <int NM>
__device__ inline void load_kf(const long long* __restrict__ kfix, int na, long long* kf) {
    for (int k = threadIdx.x; k < NM * na; k += blockDim.x) kf[(k % na) * kf_stride<NM>() + k / na] = kfix[k];
}

// LDS regions of the tile update (tile_lds_bytes).
template <typename T, int NM>
struct TileLds {
    long long* acc;
    long long* kf;
    int* wtab;
    T* wrec;
    __device__ TileLds(unsigned char* base, int ws, int na, int naxes) {
        acc = reinterpret_cast<long long*>(base);
        kf = reinterpret_cast<long long*>(base + (((size_t)ws * (NM + 1) * 8 + 15) / 16 * 16));
        wtab = reinterpret_cast<int*>(kf + (size_t)na * kf_stride<NM>());
        wrec = reinterpret_cast<T*>(reinterpret_cast<unsigned char*>(wtab) + ((size_t)(naxes + 2 * NDC) * 4 + 15) / 16 * 16);
    }
};

// One system: workgroup `blockIdx.x` moves and bins the particles of one
// tile (tile_for_block).  No occupancy request: asking the FP64 build for
// three waves per SIMD fits it into 168 VGPRs only with 14 spilled to
// scratch, 2% slower than two waves at 194 VGPRs without spills (2^24
// particles, 64^3 x 16 x 16); the FP32 build runs five waves either way.
template <typename T, int NM, bool MOVE, int NZ, bool KEEP, bool SCATTER, bool FULL = true>
__global__ void __launch_bounds__(256) k_tile_update(TileArgs<T, NM> a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int tile = tile_for_block(a.tg, blockIdx.x);
    if (tile >= a.tg.ntiles) return;
    const TileLds<T, NM> l(smem_raw, a.tg.ws, a.g.na, a.g.nx + a.g.ny + a.g.nz);
    load_kf<NM>(a.kfix, a.g.na, l.kf);
    tile_pass<T, NM, MOVE, NZ, KEEP, SCATTER, FULL, false>(a, tile, a.toff[tile], a.toff[tile + 1], true, l.acc, l.kf,
                                                           l.wtab, l.wrec, a.step_off);
}

// Many systems in one persistent launch: work item w of the concatenated
// per-system workgroup ranges [start[j], start[j+1]) is workgroup
// w - start[j] of system j.  The grid strides over all work items.
constexpr int MAX_BATCH = 256;

// Dynamic LDS: [lds_tile bytes for the tile body][njobs + 1 ints of start]; no
// static LDS, so the 64-bit accumulators at the region's start stay 8-byte
// aligned.
template <typename T, int NM, bool MOVE, int NZ, bool SCATTER, bool FULL = true>
__global__ void __launch_bounds__(256)
    k_tile_update_batched(const TileArgs<T, NM>* __restrict__ jobs, const int* __restrict__ start, int njobs,
                          int lds_tile, uint32_t step_off) {
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
        lo = __builtin_amdgcn_readfirstlane(lo);  // uniform: the job's fields come by scalar loads
        // a reference to the job in global memory, not a local copy: the
        // copy of the whole structure ends up in scratch memory
        const TileArgs<T, NM>& a = jobs[lo];
        const int tile = tile_for_block(a.tg, w - sstart[lo]);
        if (tile < a.tg.ntiles) {
            const TileLds<T, NM> l(smem_raw, a.tg.ws, a.g.na, a.g.nx + a.g.ny + a.g.nz);
            load_kf<NM>(a.kfix, a.g.na, l.kf);
            tile_pass<T, NM, MOVE, NZ, false, SCATTER, FULL>(a, tile, a.toff[tile], a.toff[tile + 1], true, l.acc,
                                                             l.kf, l.wtab, l.wrec, step_off);
        }
    }
}

// Moments of the next step: main + inflow, fixed point -> working precision,
// into the forward-FFT buffer (imaginary part 0); clears inflow.
template <typename T, int NM>
__global__ void __launch_bounds__(256)
    k_tile_convert(const long long* __restrict__ main, long long* __restrict__ inflow, int ns, double scale,
                   double rho_scale, cplx<T>* __restrict__ spec, T* __restrict__ rho, int zz_trace) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ns) return;
#pragma unroll
    for (int q = 0; q < NM; ++q) {
        // (sigma_zz from the trace: the component is not summed, 0 here)
        const long long v =
            NM == 6 && zz_trace && q == 5 ? 0 : main[(size_t)q * ns + c] + inflow[(size_t)q * ns + c];
        spec[(size_t)q * ns + c] = cplx<T>{(T)((double)v * scale), T(0)};
        inflow[(size_t)q * ns + c] = 0;
    }
    const long long cntv = main[(size_t)NM * ns + c] + inflow[(size_t)NM * ns + c];
    inflow[(size_t)NM * ns + c] = 0;
    if (rho) rho[c] = (T)((double)cntv * rho_scale);
}

// Re-sort offsets from the direction counts of the last pass.
//
// Far particles per destination tile: the segments of the tiles whose last
// pass counted far particles are scanned again (the same window test as the
// pass: a cell outside the tile's window).
template <typename T>
__global__ void __launch_bounds__(256)
    k_far_count(State<T> s, const uint32_t* __restrict__ toff, Geo<T> g, TileGeo tg,
                const uint32_t* __restrict__ dcnt, uint32_t* __restrict__ farcnt) {
    const int t = blockIdx.x;
    if (dcnt[(size_t)t * NDC + DIR_FAR] == 0) return;
    const int tz = t % tg.nt[2], ty = (t / tg.nt[2]) % tg.nt[1], tx = t / (tg.nt[2] * tg.nt[1]);
    const int ox = tx * tg.t[0], oy = ty * tg.t[1], oz = tz * tg.t[2];
    for (uint32_t i = toff[t] + threadIdx.x; i < toff[t + 1]; i += blockDim.x) {
        const int cx = bin_x(g, at(s.x, i)), cy = bin_y(g, at(s.y, i)), cz = bin_z(g, at(s.z, i));
        if (window_coord(cx, ox, tg.t[0], tg.h[0], g.nx) < 0 || window_coord(cy, oy, tg.t[1], tg.h[1], g.ny) < 0 ||
            window_coord(cz, oz, tg.t[2], tg.h[2], g.nz) < 0)
            atomicAdd(farcnt + tile_of(tg, cx, cy, cz), 1u);
    }
}

// Tile at offset (dx, dy, dz) from tile t, periodic in tile coordinates.
__device__ inline int tile_offset(const TileGeo& tg, int t, int dx, int dy, int dz) {
    const int tz = t % tg.nt[2], ty = (t / tg.nt[2]) % tg.nt[1], tx = t / (tg.nt[2] * tg.nt[1]);
    const int x = (tx + dx + tg.nt[0]) % tg.nt[0], y = (ty + dy + tg.nt[1]) % tg.nt[1],
              z = (tz + dz + tg.nt[2]) % tg.nt[2];
    return (x * tg.nt[1] + y) * tg.nt[2] + z;
}

// Particles of each tile after the re-sort: those moving into it from each
// direction, and its far arrivals.  One wavefront per tile, lane d < 27 over
// direction d.
__global__ void __launch_bounds__(256)
    k_dir_totals(const uint32_t* __restrict__ dcnt, const uint32_t* __restrict__ farcnt, TileGeo tg,
                 uint32_t* __restrict__ tcnt) {
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= tg.ntiles) return;
    const int d = __lane_id();
    uint32_t v = 0;
    if (d < DIR_FAR) v = dcnt[(size_t)tile_offset(tg, t, 1 - d / 9, 1 - (d / 3) % 3, 1 - d % 3) * NDC + d];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (d == 0) tcnt[t] = v + farcnt[t];
}

// The new segment of tile t (from toff2[t]) holds the blocks of the particles
// arriving from direction d = 0 .. 26 (source tile t - offset(d)), then its
// far arrivals; boff[source][d] is the start of each block.  Clears the far
// placement counters.  One wavefront per tile, lane d over direction d.
__global__ void __launch_bounds__(256)
    k_dir_offsets(const uint32_t* __restrict__ dcnt, const uint32_t* __restrict__ toff2, TileGeo tg,
                  uint32_t* __restrict__ boff, uint32_t* __restrict__ farbase, uint32_t* __restrict__ farctr) {
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= tg.ntiles) return;
    const int d = __lane_id();
    size_t k = 0;
    uint32_t v = 0;
    if (d < DIR_FAR) {
        k = (size_t)tile_offset(tg, t, 1 - d / 9, 1 - (d / 3) % 3, 1 - d % 3) * NDC + d;
        v = dcnt[k];
    }
    // inclusive scan over the lanes, in direction order
    uint32_t s = v;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(s, off);
        if (d >= off) s += u;
    }
    const uint32_t o = toff2[t];
    if (d < DIR_FAR) boff[k] = o + s - v;
    if (d == DIR_FAR - 1) {
        farbase[t] = o + s;
        farctr[t] = 0u;
    }
}

// Particles per tile from the particle state (initial sort).
template <typename T>
__global__ void __launch_bounds__(256)
    k_tile_count_state(State<T> s, uint64_t n, Geo<T> g, TileGeo tg, uint32_t* __restrict__ tcnt) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t t = INVALID_KEY;
    if (i < n)
        t = (uint32_t)tile_of(tg, bin_x(g, s.x[i]), bin_y(g, s.y[i]), bin_z(g, s.z[i]));
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
    return (uint32_t)tile_of(tg, bin_x(g, a.x[i]), bin_y(g, a.y[i]), bin_z(g, a.z[i]));
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

// Plain path with the particles kept in tile order (re-sorted every few
// steps; the tile geometry without halo): one workgroup per tile counts the
// particles of its segment that lie in its own cells into LDS
// ([tile cell][angle] u32) and stores those rows of the count grid with plain
// stores, so the grid needs no clearing and no global atomics for them.  A
// particle that drifted out of its tile since the last sort is listed in its
// tile's own stretch of the list (at the tile's segment offset, an LDS counter
// per tile) for k_bin_strays.  The counts are exact whatever the order.
// COUNT (steps that re-sort): particles per tile of the current positions
// into tcnt (own particles here, drifted ones by k_bin_strays).
// (Distribution::histogram_from, distribution/mod.rs:146-160.)
constexpr int PSORT_BATCH = 4;  // particles loaded per thread before binning them

template <typename T, bool COUNT>
__global__ void __launch_bounds__(256)
    k_bin_tile(State<T> st, const uint32_t* __restrict__ toff, Geo<T> g, TileGeo tg, uint32_t* __restrict__ hist,
               uint32_t* __restrict__ strays, uint32_t* __restrict__ scnt, uint32_t* __restrict__ tcnt) {
    extern __shared__ __attribute__((aligned(16))) uint32_t cnt[];  // [t0 t1 t2][na]
    __shared__ uint32_t s_stray;
    const int na = g.na;
    const int t = blockIdx.x;
    const int t0 = tg.t[0], t1 = tg.t[1], t2 = tg.t[2], tc = t0 * t1 * t2;
    const int tz = t % tg.nt[2], ty = (t / tg.nt[2]) % tg.nt[1], tx = t / (tg.nt[2] * tg.nt[1]);
    const int ox = tx * t0, oy = ty * t1, oz = tz * t2;
    const int ex = min(t0, g.nx - ox), ey = min(t1, g.ny - oy), ez = min(t2, g.nz - oz);
    const bool v4 = (na & 3) == 0;  // rows in 16-byte pieces
    if (v4)
        for (int k = threadIdx.x; k < tc * na / 4; k += blockDim.x) reinterpret_cast<uint4*>(cnt)[k] = uint4{0, 0, 0, 0};
    else
        for (int k = threadIdx.x; k < tc * na; k += blockDim.x) cnt[k] = 0u;
    if (threadIdx.x == 0) s_stray = 0u;
    __syncthreads();
    const uint32_t i0 = toff[t], i1 = toff[t + 1];
    for (uint32_t base = i0; base < i1; base += PSORT_BATCH * blockDim.x) {
        T q[PSORT_BATCH][5];
#pragma unroll
        for (int b = 0; b < PSORT_BATCH; ++b) {
            const uint32_t i = base + b * blockDim.x + threadIdx.x;
            if (i < i1) {
                q[b][0] = at(st.x, i);
                q[b][1] = at(st.y, i);
                q[b][2] = at(st.z, i);
                q[b][3] = at(st.phi, i);
                q[b][4] = at(st.theta, i);
            }
        }
#pragma unroll
        for (int b = 0; b < PSORT_BATCH; ++b) {
            if (base + b * blockDim.x + threadIdx.x >= i1) continue;
            const int ix = bin_x(g, q[b][0]), iy = bin_y(g, q[b][1]), iz = bin_z(g, q[b][2]);
            const int a = angle_of(g, q[b][3], q[b][4]);
            const int lx = ix - ox, ly = iy - oy, lz = iz - oz;
            if (lx >= 0 && lx < ex && ly >= 0 && ly < ey && lz >= 0 && lz < ez) {
                atomicAdd(cnt + ((lx * t1 + ly) * t2 + lz) * na + a, 1u);
            } else {
                const uint32_t k = atomicAdd(&s_stray, 1u);
                strays[i0 + k] = (uint32_t)((ix * g.ny + iy) * g.nz + iz) * (uint32_t)na + (uint32_t)a;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        scnt[t] = s_stray;
        if (COUNT) atomicAdd(tcnt + t, (i1 - i0) - s_stray);
    }
    // rows of the own cells: z-runs of ez cells are contiguous in the grid
    const int w = v4 ? na / 4 : na;  // pieces per row
    const int rows = ex * ey * ez * w;
    for (int k = threadIdx.x; k < rows; k += blockDim.x) {
        const int a = k % w, c = k / w;
        const int lz = c % ez, ly = (c / ez) % ey, lx = c / (ez * ey);
        const size_t cell = (size_t)((ox + lx) * g.ny + oy + ly) * g.nz + oz + lz;
        const int lc = (lx * t1 + ly) * t2 + lz;
        if (v4)
            reinterpret_cast<uint4*>(hist)[cell * w + a] = reinterpret_cast<const uint4*>(cnt)[lc * w + a];
        else
            hist[cell * na + a] = cnt[lc * na + a];
    }
}

// The drifted particles' counts, one wavefront per tile over the tile's
// stretch of the list, after every tile stored its rows; COUNT: their new
// tile's particle count.
template <bool COUNT>
__global__ void __launch_bounds__(64)
    k_bin_strays(const uint32_t* __restrict__ strays, const uint32_t* __restrict__ toff,
                 const uint32_t* __restrict__ scnt, int ny, int nz, int na, TileGeo tg, uint32_t* __restrict__ hist,
                 uint32_t* __restrict__ tcnt) {
    const int t = blockIdx.x;
    const uint32_t i0 = toff[t], m = scnt[t];
    for (uint32_t j = threadIdx.x; j < m; j += blockDim.x) {
        const uint32_t key = strays[i0 + j];
        atomicAdd(hist + key, 1u);
        if (COUNT) {
            const uint32_t cell = key / (uint32_t)na;
            const int iz = (int)(cell % (uint32_t)nz), iy = (int)((cell / (uint32_t)nz) % (uint32_t)ny),
                      ix = (int)(cell / ((uint32_t)nz * (uint32_t)ny));
            atomicAdd(tcnt + tile_of(tg, ix, iy, iz), 1u);
        }
    }
}

// Re-sort of the tile-ordered plain path, one workgroup per tile of the
// current order: particles still in the tile fill its new segment (offsets
// toff2, from the counts of k_bin_tile / k_bin_strays with COUNT) from the
// front, one LDS reservation per wavefront; the others fill their new tile's
// segment from the back (back: zero per tile on entry).
template <typename T>
__global__ void __launch_bounds__(256)
    k_psort_permute(State<T> a, State<T> b, const uint32_t* __restrict__ toff, const uint32_t* __restrict__ toff2,
                    Geo<T> g, TileGeo tg, uint32_t* __restrict__ back) {
    __shared__ uint32_t s_front;
    const int t = blockIdx.x;
    if (threadIdx.x == 0) s_front = 0u;
    __syncthreads();
    const uint32_t i0 = toff[t], i1 = toff[t + 1], f0 = toff2[t];
    const int lane = __lane_id();
    const unsigned long long below = (1ull << lane) - 1ull;
    for (uint32_t base = i0; base < i1; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < i1;
        int d = -1;
        if (valid) d = tile_of(tg, bin_x(g, at(a.x, i)), bin_y(g, at(a.y, i)), bin_z(g, at(a.z, i)));
        const bool own = valid && d == t;
        const unsigned long long mask = __ballot(own);
        uint32_t wb = 0;
        if (lane == 0 && mask) wb = atomicAdd(&s_front, (uint32_t)__popcll(mask));
        wb = __shfl(wb, 0);
        if (own) {
            copy_particle(a, b, i, f0 + wb + (uint32_t)__popcll(mask & below));
        } else if (valid) {
            copy_particle(a, b, i, toff2[d + 1] - 1u - atomicAdd(back + d, 1u));
        }
    }
}

}  // namespace ss
