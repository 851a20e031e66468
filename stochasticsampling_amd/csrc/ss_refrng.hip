// ss_refrng.hip -- the reference's own random streams on the GPU.
//
// Noise (SS_NOISE_REFERENCE).  do_timestep draws its RandomVectors
// (src/bin/simulation/mod.rs:259-286) from T Pcg32 generators that
// Simulation::new seeds alike with seed_from_u64(seed) (mod.rs:111-128):
// chunk c of N / T + 1 particles takes generator c and draws, particle after
// particle, three StandardNormal values (rand_distr 0.2.1 ziggurat), 2 pi U
// and rayleigh_pdf(dr, U) (rand 0.7 Uniform(0, 1), mod.rs:383-385).  A
// ziggurat normal mostly takes one u64 from its generator, more after a
// rejection, so where a particle's draws begin depends on every draw before
// it in the chunk.
//
// The draws of a chunk are cut into windows of RN_WIN draws and decoded in
// parallel, like a self-synchronising code:
//   walk     one wavefront per window decodes from the window's first draw as
//            if a particle started there and records where its particles
//            start (bitmap), where the decoding leaves the window and how
//            many particles it found;
//   resolve  each window decodes again from its true entry, the exit of the
//            previous window, until it lands on a start of its own
//            speculative decoding: from there on both coincide (the decoding
//            is a function of the position), so the true particle count
//            follows from the bitmap; a window that never meets it passes its
//            own exit on;
//   check    windows whose entry changed (a predecessor did not meet its
//            speculative decoding) are resolved once more; one wavefront per
//            chunk then fixes whatever is left, in order, and scans the
//            counts into particle indices;
//   write    each window decodes from its true entry and writes its
//            particles' RandomVectors;
//   advance  each generator jumps past the draws its chunk used.
// Two decodings from different positions meet after about 550 draws (a few
// rejections), so windows of 8192 draws almost always meet inside.
//
// Inside a window one wavefront decodes 64 draws at a time: every lane makes
// its own draw (its generator jumps 128 steps per block), classifies the
// normal that would start there (length and value, LDS ring), and the scalar
// unit walks from particle start to particle start.
//
// Initial conditions: Particle::create_isotropic / create_homogeneous
// (src/particle.rs:224-297) draw five Uniform(0, 1) values per particle from
// one Pcg64Mcg; particle i starts at multiplication 5 i of the 128-bit MCG,
// reached by a jump.
//
// Rounding follows the reference: no contraction into fused multiply-adds.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ss_internal.h"

namespace ss {

constexpr uint64_t PCG_MUL = 6364136223846793005ull;
constexpr uint64_t SEED_INC = 11634580027462260723ull;  // rand_core 0.5 seed_from_u64
constexpr uint64_t MCG_HI = 0x2360ED051FC65DA4ull;        // Mcg128Xsl64 multiplier
constexpr uint64_t MCG_LO = 0x4385DF649FCCF645ull;
constexpr int RN_WIN = 8192;     // draws per window
constexpr int RN_RING = 256;     // draws kept per wavefront
constexpr int RN_WAVES = 4;      // wavefronts per workgroup
constexpr double RN_TWOPI = 2. * 3.141592653589793;
constexpr double RN_PIHALF = 3.141592653589793 / 2.;

// ---------------------------------------------------------------------------
// Generators
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t xsh_rr(uint64_t s) {
    const uint32_t x = (uint32_t)(((s >> 18) ^ s) >> 27);
    const uint32_t r = (uint32_t)(s >> 59);
    return (x >> r) | (x << ((32u - r) & 31u));
}

// LCG jump (Brown, "Random number generation with arbitrary strides", 1994):
// delta steps of s -> PCG_MUL s + inc equal s -> mult s + plus.
__host__ __device__ inline void lcg_jump(uint64_t delta, uint64_t inc, uint64_t& mult, uint64_t& plus) {
    uint64_t cm = PCG_MUL, cp = inc, am = 1, ap = 0;
    while (delta) {
        if (delta & 1) {
            am *= cm;
            ap = ap * cm + cp;
        }
        cp = (cm + 1) * cp;
        cm *= cm;
        delta >>= 1;
    }
    mult = am;
    plus = ap;
}

// One u64 of Pcg32 (next_u64_via_u32: low word first) from state s.
__host__ __device__ inline uint64_t pcg_u64(uint64_t s, uint64_t inc) {
    const uint32_t lo = xsh_rr(s);
    const uint32_t hi = xsh_rr(s * PCG_MUL + inc);
    return ((uint64_t)hi << 32) | lo;
}

// Bits 12..63 as the mantissa of a double in [2^e, 2^(e+1)).
__device__ inline double mant_exp(uint64_t bits, int e) {
    return __longlong_as_double((long long)((bits >> 12) | ((uint64_t)(1023 + e) << 52)));
}
// rand 0.7: Uniform(0, 1) (scale 1, low 0), Standard f64, Open01 f64.
__device__ inline double rn_uniform01(uint64_t b) { return (mant_exp(b, 0) - 1.0) * 1.0 + 0.0; }
__device__ inline double rn_gen_f64(uint64_t b) { return (double)(b >> 11) * (1.0 / 9007199254740992.0); }
__device__ inline double rn_open01(uint64_t b) { return mant_exp(b, 0) - (1.0 - 2.220446049250313e-16 / 2.0); }

// ---------------------------------------------------------------------------
// Per-chunk and per-window records
// ---------------------------------------------------------------------------
struct RnChunk {
    uint64_t state, inc;  // generator at the step's first draw
    uint64_t begin;       // first particle
    uint64_t m;           // particles
    uint64_t w0, nw;      // windows
    uint64_t kend;        // draws used this step
    uint64_t pad;
};

struct RnWin {
    uint64_t sexit;   // exit of the speculative decoding
    uint64_t tentry;  // entry of the last resolution (~0: none)
    uint64_t texit;   // exit of the decoding from tentry (when it did not meet)
    uint64_t base;    // chunk index of the window's first particle
    uint64_t input;   // entry for a repeated resolution
    uint32_t scnt;    // particles of the speculative decoding
    uint32_t tcnt;    // particles of the true decoding
    uint32_t merged;  // the decoding from tentry met the speculative one
    uint32_t redo;
};

struct RnTab {
    double x[257], f[257];
};

// Ziggurat tables of rand_distr's normal: x_0 = V / f(R), x_1 = R,
// x_i = f^-1(V / x_(i-1) + f(x_(i-1))), x_256 = 0, f_i = f(x_i),
// f(x) = exp(-x^2 / 2), R = 3.6541528853610088, V = 0.00492867323399.  The
// crate holds these numbers printed with 18 decimals; printing and reading
// back gives the same doubles.
static void ziggurat_tables(RnTab& t) {
    const double R = 3.6541528853610088, V = 0.00492867323399;
    double xv[257], fv[257];
    xv[0] = V / std::exp(-R * R / 2.0);
    xv[1] = R;
    for (int i = 2; i < 256; ++i) {
        const double last = xv[i - 1];
        xv[i] = std::sqrt(-2.0 * std::log(V / last + std::exp(-last * last / 2.0)));
    }
    xv[256] = 0.0;
    for (int i = 0; i < 257; ++i) fv[i] = std::exp(-xv[i] * xv[i] / 2.0);
    char buf[64];
    for (int i = 0; i < 257; ++i) {
        snprintf(buf, sizeof buf, "%.18f", xv[i]);
        t.x[i] = strtod(buf, nullptr);
        snprintf(buf, sizeof buf, "%.18f", fv[i]);
        t.f[i] = strtod(buf, nullptr);
    }
}

// ---------------------------------------------------------------------------
// Wavefront decoder
// ---------------------------------------------------------------------------
struct RnRing {
    uint64_t* d;  // draw at position p in slot p % RN_RING
    uint8_t* nl;  // draws of the normal starting at p (0: not classified)
    double* nv;   // its value
};

struct RnGen {
    uint64_t st;          // per lane: state before this lane's draw of block g
    uint64_t a128, c128;  // 128 generator steps (one block)
    uint64_t g;           // next block to generate (wave-uniform)
    uint64_t s0, inc;     // the chunk's generator at draw 0
};

__device__ inline uint64_t rn_far(const RnGen& G, uint64_t pos) {
    uint64_t m, p;
    lcg_jump(2 * pos, G.inc, m, p);
    return pcg_u64(m * G.s0 + p, G.inc);
}

__device__ inline uint64_t rn_draw(const RnGen& G, const RnRing& R, uint64_t pos) {
    return (pos < G.g && pos + RN_RING >= G.g) ? R.d[pos % RN_RING] : rn_far(G, pos);
}

// StandardNormal (rand_distr 0.2.1) starting at draw pos: draws used and value.
// The rejection loops end with probability one; RN_MAX_DRAWS bounds them so
// that every wavefront finishes even on input no generator produces (a normal
// needs more than 40 draws with probability below 1e-40).
constexpr uint32_t RN_MAX_DRAWS = 4096;

__device__ inline void rn_normal(const RnGen& G, const RnRing& R, const RnTab& T, uint64_t pos, uint32_t& len,
                                 double& val) {
    uint32_t k = 0;
    while (k < RN_MAX_DRAWS) {
        const uint64_t b = rn_draw(G, R, pos + k);
        const int i = (int)(b & 0xff);
        const double u = mant_exp(b, 1) - 3.0;
        const double x = u * T.x[i];
        if (fabs(x) < T.x[i + 1]) {
            len = k + 1;
            val = x;
            return;
        }
        if (i == 0) {  // zero_case: the tail beyond R from pairs of Open01 draws
            uint32_t q = k + 1;
            double xx, yy;
            do {
                xx = log(rn_open01(rn_draw(G, R, pos + q))) / T.x[1];
                yy = log(rn_open01(rn_draw(G, R, pos + q + 1)));
                q += 2;
            } while (-2.0 * yy < xx * xx && q < RN_MAX_DRAWS);
            len = q;
            val = u < 0.0 ? xx - T.x[1] : T.x[1] - xx;
            return;
        }
        const double g = rn_gen_f64(rn_draw(G, R, pos + k + 1));
        const double lhs = T.f[i + 1] + (T.f[i] - T.f[i + 1]) * g;
        const double arg = -x * x / 2.0;  // in [-6.7, 0]: x is below x_1 here
        // exp(arg) in single precision (relative error below 2e-6) decides
        // unless lhs lies that close to it; then the double exp does
        const float ef = __expf((float)arg);
        const double dd = lhs - (double)ef;
        const bool accept = fabs(dd) > 1e-5 * (double)ef ? dd < 0.0 : lhs < exp(arg);
        if (accept) {
            len = k + 2;
            val = x;
            return;
        }
        k += 2;
    }
    len = k;
    val = 0.0;
}

__device__ inline void rn_gen_block(RnGen& G, RnRing& R, int lane) {
    R.d[(G.g + lane) % RN_RING] = pcg_u64(G.st, G.inc);
    G.st = G.a128 * G.st + G.c128;
    G.g += 64;
}

template <bool VAL>
__device__ inline void rn_class_block(const RnGen& G, RnRing& R, const RnTab& T, uint64_t b, int lane) {
    uint32_t len;
    double v;
    rn_normal(G, R, T, b + lane, len, v);
    R.nl[(b + lane) % RN_RING] = len > 255 ? 0 : (uint8_t)len;
    if (VAL) R.nv[(b + lane) % RN_RING] = v;
}

template <bool VAL>
__device__ inline void rn_len_at(const RnGen& G, const RnRing& R, const RnTab& T, uint64_t pos, uint64_t cls_end,
                                 uint32_t& len, double& val) {
    uint32_t l = pos < cls_end ? R.nl[pos % RN_RING] : 0u;
    if (l) {
        len = l;
        if (VAL) val = R.nv[pos % RN_RING];
    } else {
        rn_normal(G, R, T, pos, len, val);
    }
}

enum { RN_WALK = 0, RN_RESOLVE = 1, RN_WRITE = 2 };

struct RnOut {
    void* noise;  // n records of 8 values of T
    const uint32_t* inv;  // slot of each particle id (null: the id)
    uint64_t n;
    double dr;
    int fp64;
};

struct RnResult {
    uint64_t s;      // first particle start at or after stop (or where it stopped)
    uint64_t cnt;    // particles decoded before stopping
    int merged;      // RESOLVE: met the speculative decoding at s
    uint64_t kend;   // WRITE: draw after the chunk's last particle (~0 if not reached)
};

// One particle's RandomVector as a record of 8 values (the update reads it
// with 16-byte loads).
__device__ inline void rn_store(const RnOut& o, uint64_t slot, const double v[5]) {
    if (o.fp64) {
        double2* p = reinterpret_cast<double2*>((double*)o.noise + slot * 8);
        p[0] = make_double2(v[0], v[1]);
        p[1] = make_double2(v[2], v[3]);
        p[2] = make_double2(v[4], 0.0);
    } else {
        float4* p = reinterpret_cast<float4*>((float*)o.noise + slot * 8);
        p[0] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
        p[1] = make_float4((float)v[4], 0.f, 0.f, 0.f);
    }
}

// Decode particles from draw `start` while they start before `stop`.
// WALK: set their bits in `bitmap` (window at wbase).  RESOLVE: stop at the
// first start whose bit is set.  WRITE: write particles base, base+1, ...
// below m of the chunk whose first particle is pbegin.
// Particle starts of one block from entry e (offset of the first start in
// the block; 64 or more: none): lane k < 16 gets the offset of start k, or 64
// once the chain has left the block.  rl: length of the particle that would
// start at each lane's draw.  Pointer doubling over the 64 lanes: j_b jumps
// 2^b particles, and lane k composes the jumps of the bits of k.
__device__ inline uint32_t rn_chain(uint32_t rl, uint32_t e, int lane) {
    const uint32_t j0 = min((uint32_t)lane + rl, 64u);
    uint32_t t = (uint32_t)__shfl((int)j0, (int)(j0 & 63));
    const uint32_t j1 = j0 >= 64 ? 64u : t;
    t = (uint32_t)__shfl((int)j1, (int)(j1 & 63));
    const uint32_t j2 = j1 >= 64 ? 64u : t;
    t = (uint32_t)__shfl((int)j2, (int)(j2 & 63));
    const uint32_t j3 = j2 >= 64 ? 64u : t;
    uint32_t p = min(e, 64u);
    t = (uint32_t)__shfl((int)j0, (int)(p & 63));
    if (lane & 1) p = p >= 64 ? 64u : t;
    t = (uint32_t)__shfl((int)j1, (int)(p & 63));
    if (lane & 2) p = p >= 64 ? 64u : t;
    t = (uint32_t)__shfl((int)j2, (int)(p & 63));
    if (lane & 4) p = p >= 64 ? 64u : t;
    t = (uint32_t)__shfl((int)j3, (int)(p & 63));
    if (lane & 8) p = p >= 64 ? 64u : t;
    return p;
}

// Write queue of a wavefront (WRITE): particles wait here until 64 are ready,
// so that the rotate angle's log runs on full wavefronts and the records of
// consecutive particles are stored together.
constexpr uint32_t RN_QCAP = 80;  // 63 waiting + at most 13 starts per block
struct RnQueue {
    uint32_t* pid;  // [RN_QCAP]
    double* v;      // [4][RN_QCAP]: three normals and the axis angle
    double* u;      // [RN_QCAP]: uniform of the rotate angle
};

__device__ inline void rn_flush(const RnQueue& Q, uint32_t nq, int lane, const RnOut& out) {
    if ((uint32_t)lane < nq) {
        double v[5];
        v[0] = Q.v[0 * RN_QCAP + lane];
        v[1] = Q.v[1 * RN_QCAP + lane];
        v[2] = Q.v[2 * RN_QCAP + lane];
        v[3] = Q.v[3 * RN_QCAP + lane];
        v[4] = out.dr * sqrt(-2. * log(1. - Q.u[lane]));  // rayleigh_pdf (mod.rs:383-385)
        const uint32_t pid = Q.pid[lane];
        rn_store(out, out.inv ? out.inv[pid] : pid, v);
    }
}

// Decode particles from draw `start` while they start before `stop` (a
// block boundary, or ~0).  WALK: record their starts in `bitmap` (window at
// wbase).  RESOLVE: stop at the first start whose bit is set.  WRITE: write
// particles base, base+1, ... below m of the chunk whose first particle is
// pbegin.
template <int MODE>
__device__ RnResult rn_decode(RnGen& G, RnRing& R, const RnQueue& Q, const RnTab& T, uint64_t start, uint64_t stop,
                              int lane, uint32_t* bitmap, uint64_t wbase, uint64_t base, uint64_t m, uint64_t pbegin,
                              const RnOut& out) {
    constexpr bool VAL = MODE == RN_WRITE;
    RnResult res;
    res.merged = 0;
    res.kend = ~0ull;
    res.s = start;
    res.cnt = 0;
    if (start >= stop) return res;
    uint64_t P = start & ~63ull;
    // lane generators at block P
    {
        uint64_t mm, pp;
        lcg_jump(2 * (P + (uint64_t)lane), G.inc, mm, pp);
        G.st = mm * G.s0 + pp;
        G.g = P;
    }
    rn_gen_block(G, R, lane);
    rn_gen_block(G, R, lane);
    rn_gen_block(G, R, lane);
    __builtin_amdgcn_wave_barrier();
    rn_class_block<VAL>(G, R, T, P, lane);
    rn_class_block<VAL>(G, R, T, P + 64, lane);
    uint64_t cls_end = P + 128;
    __builtin_amdgcn_wave_barrier();
    uint32_t e = (uint32_t)(start - P);
    uint64_t cnt = 0;
    uint32_t qn = 0;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;  // lanes below this one
    for (;;) {
        // length of the particle that would start at this lane's draw
        const uint64_t q = P + (uint64_t)lane;
        uint32_t n1, n2, n3;
        double v1 = 0.0, v2 = 0.0, v3 = 0.0;
        rn_len_at<VAL>(G, R, T, q, cls_end, n1, v1);
        rn_len_at<VAL>(G, R, T, q + n1, cls_end, n2, v2);
        rn_len_at<VAL>(G, R, T, q + n1 + n2, cls_end, n3, v3);
        const uint32_t rl = n1 + n2 + n3 + 2;
        const uint32_t pk = rn_chain(rl, e, lane);
        uint32_t count = (uint32_t)__popcll(__ballot(lane < 16 && pk < 64));
        uint32_t enext;
        if (count == 0) {
            enext = e - 64;
        } else {
            const uint32_t last = (uint32_t)__builtin_amdgcn_readlane((int)pk, (int)count - 1);
            enext = last + (uint32_t)__builtin_amdgcn_readlane((int)rl, (int)last) - 64;
        }
        bool done = false;
        if (MODE == RN_RESOLVE) {
            const uint64_t o = (P - wbase) / 32;
            const uint64_t bm = (uint64_t)bitmap[o] | ((uint64_t)bitmap[o + 1] << 32);
            const uint64_t hit = __ballot((uint32_t)lane < count && ((bm >> (pk & 63)) & 1ull));
            if (hit) {
                const uint32_t k = (uint32_t)__ffsll((unsigned long long)hit) - 1;
                res.merged = 1;
                res.s = P + (uint32_t)__builtin_amdgcn_readlane((int)pk, (int)k);
                res.cnt = cnt + k;
                return res;
            }
        }
        if (MODE == RN_WRITE && base + cnt + count >= m) {
            const uint32_t kl = (uint32_t)(m - 1 - (base + cnt));
            const uint32_t pl = (uint32_t)__builtin_amdgcn_readlane((int)pk, (int)kl);
            res.kend = P + pl + (uint32_t)__builtin_amdgcn_readlane((int)rl, (int)pl);
            count = kl + 1;
            done = true;
        }
        // this block's starts as a bit mask
        uint64_t bits = (uint32_t)lane < count ? (1ull << (pk & 63)) : 0ull;
        for (int off = 8; off > 0; off >>= 1) bits |= __shfl_xor(bits, off);
        const uint64_t mask = __shfl(bits, 0);
        if (MODE == RN_WALK && lane == 0) {
            const uint64_t o = (P - wbase) / 32;
            bitmap[o] = (uint32_t)mask;
            bitmap[o + 1] = (uint32_t)(mask >> 32);
        }
        if (MODE == RN_WRITE) {
            if ((mask >> lane) & 1ull) {
                const uint32_t slot = qn + (uint32_t)__popcll(mask & lt);
                const uint64_t up = q + n1 + n2 + n3;
                Q.pid[slot] = (uint32_t)(pbegin + base + cnt + (slot - qn));
                Q.v[0 * RN_QCAP + slot] = v1;
                Q.v[1 * RN_QCAP + slot] = v2;
                Q.v[2 * RN_QCAP + slot] = v3;
                Q.v[3 * RN_QCAP + slot] = RN_TWOPI * rn_uniform01(rn_draw(G, R, up));
                Q.u[slot] = rn_uniform01(rn_draw(G, R, up + 1));
            }
            qn += count;
            __builtin_amdgcn_wave_barrier();
            if (qn >= 64) {
                rn_flush(Q, 64, lane, out);
                if ((uint32_t)lane < qn - 64) {
                    Q.pid[lane] = Q.pid[64 + lane];
                    for (int k = 0; k < 4; ++k) Q.v[k * RN_QCAP + lane] = Q.v[k * RN_QCAP + 64 + lane];
                    Q.u[lane] = Q.u[64 + lane];
                }
                qn -= 64;
                __builtin_amdgcn_wave_barrier();
            }
        }
        cnt += count;
        res.s = P + 64 + enext;
        if (done || P + 64 >= stop) break;
        // next block: generate three blocks ahead, classify two ahead
        P += 64;
        e = enext;
        __builtin_amdgcn_wave_barrier();
        rn_gen_block(G, R, lane);
        __builtin_amdgcn_wave_barrier();
        rn_class_block<VAL>(G, R, T, P + 64, lane);
        cls_end = P + 128;
        __builtin_amdgcn_wave_barrier();
    }
    if (MODE == RN_WRITE && qn) rn_flush(Q, qn, lane, out);
    res.cnt = cnt;
    return res;
}

// Set bits of the window bitmap at or after bit o (RN_WIN bits).
__device__ inline uint32_t rn_popcount_from(const uint32_t* bitmap, uint64_t o, int lane) {
    uint32_t c = 0;
    for (int w = lane; w < RN_WIN / 32; w += 64) {
        uint32_t v = bitmap[w];
        const uint64_t lo = (uint64_t)w * 32;
        if (lo + 32 <= o) v = 0;
        else if (lo < o) v &= ~0u << (o - lo);
        c += (uint32_t)__popc(v);
    }
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
    return c;
}

template <bool WRITE>
struct RnShared {
    RnTab tab;
    uint64_t d[RN_WAVES][RN_RING];
    uint8_t nl[RN_WAVES][RN_RING];
    double nv[RN_WAVES][WRITE ? RN_RING : 1];
    uint32_t qpid[RN_WAVES][WRITE ? RN_QCAP : 1];
    double qv[RN_WAVES][WRITE ? 4 * RN_QCAP : 1];
    double qu[RN_WAVES][WRITE ? RN_QCAP : 1];
};

template <bool WRITE>
__device__ inline void rn_setup(RnShared<WRITE>& sh, const RnTab* tab, RnRing& R, RnQueue& Q, int wave) {
    for (int i = threadIdx.x; i < 257; i += blockDim.x) {
        sh.tab.x[i] = tab->x[i];
        sh.tab.f[i] = tab->f[i];
    }
    R.d = sh.d[wave];
    R.nl = sh.nl[wave];
    R.nv = sh.nv[wave];
    Q.pid = sh.qpid[wave];
    Q.v = sh.qv[wave];
    Q.u = sh.qu[wave];
    __syncthreads();
}

__device__ inline void rn_gen_init(RnGen& G, const RnChunk& c) {
    G.s0 = c.state;
    G.inc = c.inc;
    lcg_jump(128, c.inc, G.a128, G.c128);
}

// Resolution of window j of chunk c from entry t; lane 0 stores the result.
__device__ inline void rn_resolve_one(RnGen& G, RnRing& R, const RnQueue& Q, const RnTab& T, const RnChunk& c,
                                      RnWin* win, const uint32_t* bitmap, uint64_t w, uint64_t t, int lane) {
    const uint64_t j = w - c.w0;
    const uint64_t wbase = j * RN_WIN, stop = wbase + RN_WIN;
    const uint32_t* bm = bitmap + w * (RN_WIN / 32);
    RnOut none{nullptr, nullptr, 0, 0.0, 0};
    RnResult r = rn_decode<RN_RESOLVE>(G, R, Q, T, t, stop, lane, const_cast<uint32_t*>(bm), wbase, 0, 0, 0, none);
    uint32_t tcnt = (uint32_t)r.cnt;
    if (r.merged) tcnt += rn_popcount_from(bm, r.s - wbase, lane);
    if (lane == 0) {
        win[w].tentry = t;
        win[w].merged = (uint32_t)r.merged;
        win[w].texit = r.merged ? win[w].sexit : r.s;
        win[w].tcnt = tcnt;
    }
}

__device__ inline uint64_t rn_expected(const RnWin* win, uint64_t w) {
    return win[w - 1].merged ? win[w - 1].sexit : win[w - 1].texit;
}

__global__ void __launch_bounds__(64 * RN_WAVES)
    k_rn_walk(const RnChunk* __restrict__ ch, RnWin* __restrict__ win, const uint32_t* __restrict__ win_chunk,
              uint32_t* __restrict__ bitmap, const RnTab* __restrict__ tab, uint64_t nwin) {
    __shared__ RnShared<false> sh;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    RnRing R;
    RnQueue Q;
    rn_setup(sh, tab, R, Q, wave);
    const uint64_t w = (uint64_t)blockIdx.x * RN_WAVES + wave;
    if (w >= nwin) return;
    const RnChunk c = ch[win_chunk[w]];
    RnGen G;
    rn_gen_init(G, c);
    const uint64_t j = w - c.w0, wbase = j * RN_WIN;
    RnOut none{nullptr, nullptr, 0, 0.0, 0};
    RnResult r = rn_decode<RN_WALK>(G, R, Q, sh.tab, wbase, wbase + RN_WIN, lane, bitmap + w * (RN_WIN / 32), wbase,
                                    0, 0, 0, none);
    if (lane == 0) {
        win[w].sexit = r.s;
        win[w].scnt = (uint32_t)r.cnt;
        win[w].tentry = ~0ull;
        win[w].redo = 1;
        win[w].input = 0;
    }
}

// first: every window from the exit of its predecessor's speculative
// decoding; later rounds: the windows marked by k_rn_check.
__global__ void __launch_bounds__(64 * RN_WAVES)
    k_rn_resolve(const RnChunk* __restrict__ ch, RnWin* __restrict__ win, const uint32_t* __restrict__ win_chunk,
                 const uint32_t* __restrict__ bitmap, const RnTab* __restrict__ tab, uint64_t nwin, int first) {
    __shared__ RnShared<false> sh;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    RnRing R;
    RnQueue Q;
    rn_setup(sh, tab, R, Q, wave);
    const uint64_t w = (uint64_t)blockIdx.x * RN_WAVES + wave;
    if (w >= nwin) return;
    const RnChunk c = ch[win_chunk[w]];
    uint64_t t;
    if (first) {
        t = w == c.w0 ? 0 : win[w - 1].sexit;
    } else {
        if (!win[w].redo) return;
        t = win[w].input;
    }
    RnGen G;
    rn_gen_init(G, c);
    rn_resolve_one(G, R, Q, sh.tab, c, win, bitmap, w, t, lane);
}

__global__ void __launch_bounds__(256)
    k_rn_check(const RnChunk* __restrict__ ch, RnWin* __restrict__ win, const uint32_t* __restrict__ win_chunk,
               uint64_t nwin) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwin) return;
    const RnChunk c = ch[win_chunk[w]];
    uint32_t redo = 0;
    uint64_t in = 0;
    if (w != c.w0) {
        in = rn_expected(win, w);
        redo = win[w].tentry != in;
    }
    win[w].redo = redo;
    win[w].input = in;
}

// One wavefront per chunk: resolve, in order, every window whose entry is not
// its predecessor's true exit; number the particles; decode past the last
// window if the windows held fewer than m particles.
__global__ void __launch_bounds__(64 * RN_WAVES)
    k_rn_finish(RnChunk* __restrict__ ch, RnWin* __restrict__ win, const uint32_t* __restrict__ bitmap,
                const RnTab* __restrict__ tab, int nchunks, RnOut out) {
    __shared__ RnShared<true> sh;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    RnRing R;
    RnQueue Q;
    rn_setup(sh, tab, R, Q, wave);
    const int cidx = blockIdx.x * RN_WAVES + wave;
    if (cidx >= nchunks) return;
    const RnChunk c = ch[cidx];
    RnGen G;
    rn_gen_init(G, c);
    // lane 0 stores the results of a resolution; the other lanes read them
    const volatile RnWin* vw = win;
    uint64_t jlo = 1;
    for (;;) {
        uint64_t bad = ~0ull;
        for (uint64_t j0 = jlo; j0 < c.nw && bad == ~0ull; j0 += 64) {
            const uint64_t j = j0 + lane;
            bool b = false;
            if (j < c.nw) {
                const uint64_t w = c.w0 + j;
                const uint64_t exp_in = vw[w - 1].merged ? vw[w - 1].sexit : vw[w - 1].texit;
                b = vw[w].tentry != exp_in;
            }
            const uint64_t bal = __ballot(b);
            if (bal) bad = j0 + (uint64_t)__ffsll((unsigned long long)bal) - 1;
        }
        if (bad == ~0ull) break;
        const uint64_t w = c.w0 + bad;
        const uint64_t t = vw[w - 1].merged ? vw[w - 1].sexit : vw[w - 1].texit;
        __builtin_amdgcn_wave_barrier();
        rn_resolve_one(G, R, Q, sh.tab, c, win, bitmap, w, t, lane);
        __threadfence_block();
        jlo = bad + 1;
    }
    // particle indices: exclusive scan of the true counts
    uint64_t run = 0;
    for (uint64_t j0 = 0; j0 < c.nw; j0 += 64) {
        const uint64_t j = j0 + lane;
        uint64_t v = j < c.nw ? vw[c.w0 + j].tcnt : 0;
        uint64_t incl = v;
        for (int off = 1; off < 64; off <<= 1) {
            const uint64_t o = __shfl_up(incl, off);
            if (lane >= off) incl += o;
        }
        if (j < c.nw) win[c.w0 + j].base = run + incl - v;
        run += __shfl(incl, 63);
    }
    uint64_t kend = ~0ull;
    if (run < c.m) {  // more particles than the windows cover: decode on
        const uint64_t w = c.w0 + c.nw - 1;
        const uint64_t t = vw[w].merged ? vw[w].sexit : vw[w].texit;
        RnResult r = rn_decode<RN_WRITE>(G, R, Q, sh.tab, t, ~0ull, lane, nullptr, 0, run, c.m, c.begin, out);
        kend = r.kend;
    }
    if (lane == 0) ch[cidx].kend = kend;
}

__global__ void __launch_bounds__(64 * RN_WAVES)
    k_rn_write(RnChunk* __restrict__ ch, const RnWin* __restrict__ win, const uint32_t* __restrict__ win_chunk,
               const RnTab* __restrict__ tab, uint64_t nwin, RnOut out) {
    __shared__ RnShared<true> sh;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    RnRing R;
    RnQueue Q;
    rn_setup(sh, tab, R, Q, wave);
    const uint64_t w = (uint64_t)blockIdx.x * RN_WAVES + wave;
    if (w >= nwin) return;
    const uint32_t ci = win_chunk[w];
    const RnChunk c = ch[ci];
    if (win[w].base >= c.m || win[w].tcnt == 0) return;
    RnGen G;
    rn_gen_init(G, c);
    const uint64_t j = w - c.w0, stop = (j + 1) * RN_WIN;
    RnResult r = rn_decode<RN_WRITE>(G, R, Q, sh.tab, win[w].tentry, stop, lane, nullptr, 0, win[w].base, c.m,
                                     c.begin, out);
    if (r.kend != ~0ull && lane == 0) ch[ci].kend = r.kend;
}

__global__ void k_rn_advance(RnChunk* __restrict__ ch, int nchunks) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nchunks) return;
    uint64_t m, p;
    lcg_jump(2 * ch[c].kend, ch[c].inc, m, p);
    ch[c].state = m * ch[c].state + p;
}

// ---------------------------------------------------------------------------
// Host side of SS_NOISE_REFERENCE
// ---------------------------------------------------------------------------
struct RefNoise {
    int streams = 0;          // generators (T)
    int nchunks = 0;          // chunks with particles
    uint64_t n = 0;
    uint64_t nwin = 0;
    std::vector<uint64_t> host_state;  // generators of chunks without particles (never advanced)
    RnChunk* ch = nullptr;
    RnWin* win = nullptr;
    uint32_t* win_chunk = nullptr;
    uint32_t* bitmap = nullptr;
    RnTab* tab = nullptr;
};

// Pcg32::seed_from_u64 (rand_core 0.5 + rand_pcg 0.2 from_seed / from_state_incr).
static void pcg32_seed_from_u64(uint64_t s, uint64_t& state, uint64_t& inc) {
    uint32_t w[4];
    for (int k = 0; k < 4; ++k) {
        s = s * PCG_MUL + SEED_INC;
        w[k] = xsh_rr(s);
    }
    const uint64_t st = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
    inc = ((uint64_t)w[2] | ((uint64_t)w[3] << 32)) | 1u;
    state = (st + inc) * PCG_MUL + inc;
}

int refnoise_create(RefNoise** out, uint64_t n, int streams, uint64_t seed) {
    *out = nullptr;
    if (streams < 1) return fail(SS_ERR_INVALID, "No environment variable 'RAYON_NUM_THREADS' set.");
    RefNoise* r = new RefNoise();
    r->streams = streams;
    r->n = n;
    const uint64_t cs = n / (uint64_t)streams + 1;  // mod.rs:261
    r->nchunks = (int)std::min<uint64_t>((uint64_t)streams, (n + cs - 1) / cs);
    std::vector<RnChunk> hc(r->nchunks);
    std::vector<uint32_t> wc;
    uint64_t st, inc;
    pcg32_seed_from_u64(seed, st, inc);
    r->host_state.assign(2 * (size_t)streams, 0);
    for (int c = 0; c < streams; ++c) {
        r->host_state[2 * c] = st;
        r->host_state[2 * c + 1] = inc;
    }
    for (int c = 0; c < r->nchunks; ++c) {
        RnChunk& k = hc[c];
        k.state = st;
        k.inc = inc;
        k.begin = (uint64_t)c * cs;
        k.m = std::min<uint64_t>(cs, n - k.begin);
        // about 5.07 draws per particle; the tail past the windows is decoded
        // by k_rn_finish should a chunk ever need more
        const double need = 5.08 * (double)k.m + 64.0 * std::sqrt((double)k.m) + 256.0;
        k.nw = std::max<uint64_t>(1, (uint64_t)std::ceil(need / RN_WIN));
        k.w0 = r->nwin;
        k.kend = 0;
        k.pad = 0;
        r->nwin += k.nw;
        for (uint64_t j = 0; j < k.nw; ++j) wc.push_back((uint32_t)c);
    }
    RnTab t;
    ziggurat_tables(t);
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMalloc(&r->ch, sizeof(RnChunk) * std::max(1, r->nchunks));
    if (e == hipSuccess) e = hipMalloc(&r->win, sizeof(RnWin) * std::max<uint64_t>(1, r->nwin));
    if (e == hipSuccess) e = hipMalloc(&r->win_chunk, sizeof(uint32_t) * std::max<uint64_t>(1, r->nwin));
    if (e == hipSuccess) e = hipMalloc(&r->bitmap, sizeof(uint32_t) * (RN_WIN / 32) * std::max<uint64_t>(1, r->nwin));
    if (e == hipSuccess) e = hipMalloc(&r->tab, sizeof(RnTab));
    if (e == hipSuccess && r->nchunks)
        e = hipMemcpy(r->ch, hc.data(), sizeof(RnChunk) * r->nchunks, hipMemcpyHostToDevice);
    if (e == hipSuccess && r->nwin)
        e = hipMemcpy(r->win_chunk, wc.data(), sizeof(uint32_t) * r->nwin, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(r->tab, &t, sizeof(RnTab), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        refnoise_destroy(r);
        return fail(SS_ERR_ALLOC, std::string("reference noise buffers: ") + hipGetErrorString(e));
    }
    *out = r;
    return SS_OK;
}

void refnoise_destroy(RefNoise* r) {
    if (!r) return;
    (void)hipFree(r->ch);
    (void)hipFree(r->win);
    (void)hipFree(r->win_chunk);
    (void)hipFree(r->bitmap);
    (void)hipFree(r->tab);
    delete r;
}

int refnoise_streams(const RefNoise* r) { return r ? r->streams : 0; }

hipError_t refnoise_enqueue(RefNoise* r, void* noise, bool fp64, double dr, const uint32_t* inv,
                            hipStream_t stream) {
    if (r->nchunks == 0) return hipSuccess;
    const dim3 wblk(64 * RN_WAVES);
    const dim3 wgrid((unsigned)((r->nwin + RN_WAVES - 1) / RN_WAVES));
    RnOut out{noise, inv, r->n, dr, fp64 ? 1 : 0};
    hipLaunchKernelGGL(k_rn_walk, wgrid, wblk, 0, stream, r->ch, r->win, r->win_chunk, r->bitmap, r->tab, r->nwin);
    hipLaunchKernelGGL(k_rn_resolve, wgrid, wblk, 0, stream, r->ch, r->win, r->win_chunk, r->bitmap, r->tab,
                       r->nwin, 1);
    hipLaunchKernelGGL(k_rn_check, dim3((unsigned)((r->nwin + 255) / 256)), dim3(256), 0, stream, r->ch, r->win,
                       r->win_chunk, r->nwin);
    hipLaunchKernelGGL(k_rn_resolve, wgrid, wblk, 0, stream, r->ch, r->win, r->win_chunk, r->bitmap, r->tab,
                       r->nwin, 0);
    hipLaunchKernelGGL(k_rn_finish, dim3((unsigned)((r->nchunks + RN_WAVES - 1) / RN_WAVES)), wblk, 0, stream,
                       r->ch, r->win, r->bitmap, r->tab, r->nchunks, out);
    hipLaunchKernelGGL(k_rn_write, wgrid, wblk, 0, stream, r->ch, r->win, r->win_chunk, r->tab, r->nwin, out);
    hipLaunchKernelGGL(k_rn_advance, dim3((unsigned)((r->nchunks + 63) / 64)), dim3(64), 0, stream, r->ch,
                       r->nchunks);
    return hipGetLastError();
}

int refnoise_get(RefNoise* r, uint64_t* st, int count, hipStream_t stream) {
    if (count < r->streams) return fail(SS_ERR_INVALID, "rng state buffer holds fewer than noise_streams entries");
    std::vector<RnChunk> hc(r->nchunks);
    if (r->nchunks) {
        SS_HIP_CHECK(hipMemcpyAsync(hc.data(), r->ch, sizeof(RnChunk) * r->nchunks, hipMemcpyDeviceToHost, stream));
        SS_HIP_CHECK(hipStreamSynchronize(stream));
    }
    for (int c = 0; c < r->streams; ++c) {
        st[2 * c] = c < r->nchunks ? hc[c].state : r->host_state[2 * c];
        st[2 * c + 1] = c < r->nchunks ? hc[c].inc : r->host_state[2 * c + 1];
    }
    return r->streams;
}

int refnoise_set(RefNoise* r, const uint64_t* st, int count, hipStream_t stream) {
    std::vector<RnChunk> hc(r->nchunks);
    if (r->nchunks) {
        SS_HIP_CHECK(hipMemcpyAsync(hc.data(), r->ch, sizeof(RnChunk) * r->nchunks, hipMemcpyDeviceToHost, stream));
        SS_HIP_CHECK(hipStreamSynchronize(stream));
    }
    for (int c = 0; c < r->streams && c < count; ++c) {
        if (c < r->nchunks) {
            hc[c].state = st[2 * c];
            hc[c].inc = st[2 * c + 1];
        }
        r->host_state[2 * c] = st[2 * c];
        r->host_state[2 * c + 1] = st[2 * c + 1];
    }
    if (r->nchunks) {
        SS_HIP_CHECK(hipMemcpyAsync(r->ch, hc.data(), sizeof(RnChunk) * r->nchunks, hipMemcpyHostToDevice, stream));
        SS_HIP_CHECK(hipStreamSynchronize(stream));
    }
    return SS_OK;
}

// ---------------------------------------------------------------------------
// Initial conditions (Pcg64Mcg)
// ---------------------------------------------------------------------------
struct U128 {
    uint64_t lo, hi;
};

__host__ __device__ inline U128 mul128(U128 a, U128 b) {
    U128 r;
    r.lo = a.lo * b.lo;
#ifdef __HIP_DEVICE_COMPILE__
    r.hi = __umul64hi(a.lo, b.lo) + a.lo * b.hi + a.hi * b.lo;
#else
    r.hi = (uint64_t)(((unsigned __int128)a.lo * b.lo) >> 64) + a.lo * b.hi + a.hi * b.lo;
#endif
    return r;
}

// XSL-RR output of an MCG state.
__device__ inline uint64_t xsl_rr(U128 s) {
    const uint64_t x = s.hi ^ s.lo;
    const unsigned rot = (unsigned)(s.hi >> 58);
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

// quaternion 0.3.1 mul: (a0 b0 - a.b, b a0 + a b0 + a x b).
__device__ inline void rn_qmul(double a0, const double* a, double b0, const double* b, double& r0, double* r) {
    r0 = a0 * b0 - (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
    const double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    for (int i = 0; i < 3; ++i) r[i] = (b[i] * a0 + a[i] * b0) + c[i];
}

// Particle i takes multiplications 5 i + 1 .. 5 i + 5 of the MCG; pw[k] is
// the multiplier to the power 2^k.
template <bool HOMOG>
__global__ void __launch_bounds__(256) k_ic(uint64_t n, U128 s0, const U128* __restrict__ pw, double bx, double by,
                                            double bz, double kappa, double* __restrict__ out,
                                            int* __restrict__ nan_flag) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    U128 s = s0;
    uint64_t e = 5 * i;
    for (int k = 0; e; ++k, e >>= 1)
        if (e & 1) s = mul128(s, pw[k]);
    const U128 M{MCG_LO, MCG_HI};
    double u[5];
    for (int k = 0; k < 5; ++k) {
        s = mul128(s, M);
        u[k] = rn_uniform01(xsl_rr(s));
    }
    double in[5];
    in[0] = bx * u[0];
    in[1] = by * u[1];
    in[2] = bz * u[2];
    in[3] = RN_TWOPI * u[3];
    if (HOMOG) {
        in[4] = acos(log(exp(kappa) - 2. * u[4] * sinh(kappa)) / kappa);  // pdf_homogeneous_fixpoint
        if (isnan(in[4])) atomicOr(nan_flag, 1);
    } else {
        in[4] = acos(1. - 2. * u[4]);  // pdf_sin(2 U)
    }
    double p[5];
    const double box[3] = {bx, by, bz};
    sanitize(in, box, p);  // Particle::new
    if (HOMOG) {
        const double half = -RN_PIHALF / 2.;
        const double sh = sin(half), q0 = cos(half);
        const double qv[3] = {1. * sh, 0. * sh, 0. * sh};
        const double qc[3] = {-qv[0], -qv[1], -qv[2]};
        const double cp = cos(p[3]), sp = sin(p[3]), ct = cos(p[4]), st = sin(p[4]);
        const double v[3] = {st * cp, st * sp, ct};
        double t0, tv[3], r0, w[3];
        rn_qmul(q0, qv, 0.0, v, t0, tv);
        rn_qmul(t0, tv, q0, qc, r0, w);
        const double rxy = sqrt(w[0] * w[0] + w[1] * w[1]);
        p[3] = atan2(w[1], w[0]);
        p[4] = RN_PIHALF - atan2(w[2], rxy);
    }
    for (int k = 0; k < 5; ++k) out[5 * i + k] = p[k];
}

static int create_ic(bool homog, uint64_t n, double kappa, const ss_box* box, uint64_t seed, int32_t device,
                     double* particles) {
    if (!box || (!particles && n)) return fail(SS_ERR_INVALID, "null argument");
    if (homog && kappa == 0.0)
        return fail(SS_ERR_INVALID, "Alignment of zero is the isotropic state. Please use that instead.");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SS_ERR_DEVICE, "no GPU");
    if (device < 0 || device >= ndev) return fail(SS_ERR_INVALID, "device out of range");
    if (n == 0) return SS_OK;
    SS_HIP_CHECK(hipSetDevice(device));
    // Pcg64Mcg::seed_from_u64: four PCG32 output words, little endian, forced odd
    uint32_t w[4];
    uint64_t s = seed;
    for (int k = 0; k < 4; ++k) {
        s = s * PCG_MUL + SEED_INC;
        w[k] = xsh_rr(s);
    }
    const U128 s0{((uint64_t)w[0] | ((uint64_t)w[1] << 32)) | 1u, (uint64_t)w[2] | ((uint64_t)w[3] << 32)};
    U128 pw[64];
    pw[0] = U128{MCG_LO, MCG_HI};
    for (int k = 1; k < 64; ++k) pw[k] = mul128(pw[k - 1], pw[k - 1]);
    U128* d_pw = nullptr;
    double* d_out = nullptr;
    int* d_flag = nullptr;
    hipError_t e = hipMalloc(&d_pw, sizeof(pw));
    if (e == hipSuccess) e = hipMalloc(&d_out, sizeof(double) * 5 * n);
    if (e == hipSuccess) e = hipMalloc(&d_flag, sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(d_pw, pw, sizeof(pw), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d_flag, 0, sizeof(int));
    if (e == hipSuccess) {
        const dim3 grid((unsigned)((n + 255) / 256));
        if (homog)
            hipLaunchKernelGGL(k_ic<true>, grid, dim3(256), 0, 0, n, s0, d_pw, box->x, box->y, box->z, kappa, d_out,
                               d_flag);
        else
            hipLaunchKernelGGL(k_ic<false>, grid, dim3(256), 0, 0, n, s0, d_pw, box->x, box->y, box->z, kappa, d_out,
                               d_flag);
        e = hipGetLastError();
    }
    int flag = 0;
    if (e == hipSuccess) e = hipMemcpy(particles, d_out, sizeof(double) * 5 * n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(d_pw);
    (void)hipFree(d_out);
    (void)hipFree(d_flag);
    if (e != hipSuccess) return fail(SS_ERR_DEVICE, std::string("initial condition: ") + hipGetErrorString(e));
    if (flag) return fail(SS_ERR_INVALID, "Caution, the alignment parameter is too high for the given precision.");
    return SS_OK;
}

}  // namespace ss

extern "C" int ss_create_isotropic(uint64_t n, const ss_box* box, uint64_t seed, int32_t device,
                                   double* particles) {
    ss::set_error("");
    return ss::create_ic(false, n, 0.0, box, seed, device, particles);
}

extern "C" int ss_create_homogeneous(uint64_t n, double kappa, const ss_box* box, uint64_t seed, int32_t device,
                                     double* particles) {
    ss::set_error("");
    return ss::create_ic(true, n, kappa, box, seed, device, particles);
}
