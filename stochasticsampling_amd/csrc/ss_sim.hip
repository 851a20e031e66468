// ss_sim.hip -- the simulation handle behind the C ABI: device memory, the
// per-step launch sequence (Simulation::do_timestep,
// src/bin/simulation/mod.rs:251-374) and the getters.
//
// One step:
//   bin      particles -> per-cell moments (count grid or sorted LDS counts)
//   fft      forward transform of the moment components
//   solve    Oseen projection, gradient, W/E (and magnetic field) in k-space
//   fft      inverse transform of the fields the update reads
//   pack     real parts -> one record per cell
//   update   gather + Langevin chain + move (+ new keys, sorted path)
//   sort     scan of the per-cell counts + scatter into cell order (sorted path)
#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <string>
#include <vector>

#include "ss_internal.h"
#include "ss_kernels.h"
#include "ss_spectral.h"
#include "ss_tiles.h"

namespace ss {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

static const double PI_D = 3.141592653589793;
static const double TWOPI_D = 2.0 * 3.141592653589793;

// Host copies of the reference's mesh helpers, evaluated in FP64 exactly as the
// reference does, then rounded to the working precision.
static void k_sampling(int n, double L, std::vector<double>& out) {
    out.resize(n);
    int b = (n % 2 == 0) ? n / 2 : n / 2 + 1;
    double step = TWOPI_D / L;  // fft_helper.rs:26-51
    for (int i = 0; i < b; ++i) out[i] = (double)i * step;
    for (int i = b; i < n; ++i) out[i] = (double)(i - n) * step;
}

static void linspace(double a, double b, int n, std::vector<double>& out) {
    out.resize(n);
    double step = n > 1 ? (b - a) / ((double)n - 1.0) : 0.0;
    for (int i = 0; i < n; ++i) out[i] = a + step * (double)i;
}

struct GW {
    double x, y, z, phi, theta;
};
static GW grid_width(const ss_params& p) {
    return {p.box_size.x / p.grid_size.x, p.box_size.y / p.grid_size.y, p.box_size.z / p.grid_size.z,
            TWOPI_D / p.grid_size.phi, PI_D / p.grid_size.theta};
}

static bool default_direction(const ss_params& p) {
    return p.field_direction[0] == 0.0 && p.field_direction[1] == 0.0 && p.field_direction[2] == 0.0;
}

// Moment kernels on the angular bin centres: stress_kernel (stress.rs:29-56,
// closure src/bin/simulation/mod.rs:100-104) in `nsig` components, then the
// orientation kernel (director.rs:69-92) when `mag`.  out: [nm][nphi][ntheta].
static void moment_kernels(const ss_params& p, int nsig, bool mag, std::vector<double>& out) {
    const int np = p.grid_size.phi, nt = p.grid_size.theta, na = np * nt;
    GW gw = grid_width(p);
    std::vector<double> ap, at;
    linspace(0. + gw.phi / 2., TWOPI_D - gw.phi / 2., np, ap);
    linspace(0. + gw.theta / 2., PI_D - gw.theta / 2., nt, at);
    double d[3] = {p.field_direction[0], p.field_direction[1], p.field_direction[2]};
    const bool ydir = default_direction(p);
    if (ydir) {
        d[0] = 0.0;
        d[1] = 1.0;
        d[2] = 0.0;
    }
    const int nm = nsig + (mag ? 3 : 0);
    out.assign((size_t)nm * na, 0.0);
    for (int i = 0; i < np; ++i)
        for (int j = 0; j < nt; ++j) {
            const double phi = ap[i], th = at[j];
            const double cp = cos(phi), sp = sin(phi), ct = cos(th), st = sin(th);
            double sa[9] = {-(1. / 3.) + cp * cp * st * st, cp * st * st * sp, ct * cp * st,
                            cp * st * st * sp, -(1. / 3.) + st * st * sp * sp, ct * st * sp,
                            ct * cp * st, ct * st * sp, -(1. / 3.) + ct * ct};
            double sm[9];
            if (ydir) {
                const double t[9] = {0., cp * st, 0., -cp * st, 0., -ct, 0., ct, 0.};
                for (int q = 0; q < 9; ++q) sm[q] = t[q] * 0.5;
            } else {
                const double n[3] = {st * cp, st * sp, ct};
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) sm[3 * a + b] = (n[a] * d[b] - d[a] * n[b]) * 0.5;
            }
            double sr[9] = {2. * ct * (cp * cp) * (st * st), ct * (st * st) * cos(2. * phi),
                            cos(2. * th) * cp * st, ct * (st * st) * cos(2. * phi),
                            st * sin(2. * th) * (sp * sp), cos(2. * th) * st * sp,
                            cos(2. * th) * cp * st, cos(2. * th) * st * sp, -2. * ct * (st * st)};
            for (int q = 0; q < 9; ++q) sr[q] *= 0.5;
            double K[9];
            for (int q = 0; q < 9; ++q)
                K[q] = p.stress_active * sa[q] + p.stress_magnetic * sm[q] + p.shape * sr[q];
            const int a = i * nt + j;
            if (nsig == 6) {
                const int qs[6] = {0, 1, 2, 4, 5, 8};
                for (int q = 0; q < 6; ++q) out[(size_t)q * na + a] = K[qs[q]];
            } else {
                for (int q = 0; q < 9; ++q) out[(size_t)q * na + a] = K[q];
            }
            if (mag) {
                // Orientation::new applies ang_pbc to the bin centre first
                double ph = fmod(phi, TWOPI_D), tt = fmod(th, TWOPI_D);
                if (ph < 0) ph += TWOPI_D;
                if (tt < 0) tt += TWOPI_D;
                if (tt > PI_D) {
                    ph = fmod(ph + PI_D, TWOPI_D);
                    tt = TWOPI_D - tt;
                }
                out[(size_t)(nsig + 0) * na + a] = sin(tt) * cos(ph);
                out[(size_t)(nsig + 1) * na + a] = sin(tt) * sin(ph);
                out[(size_t)(nsig + 2) * na + a] = cos(tt);
            }
        }
}

template <typename T>
static int upload(void** dst, const std::vector<double>& src) {
    std::vector<T> h(src.size());
    for (size_t i = 0; i < src.size(); ++i) h[i] = (T)src[i];
    SS_HIP_CHECK(hipMalloc(dst, std::max<size_t>(1, h.size()) * sizeof(T)));
    SS_HIP_CHECK(hipMemcpy(*dst, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return SS_OK;
}

static inline unsigned blocks_for(uint64_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

// Dynamic LDS beyond the default 64 KiB needs the kernel's limit raised
// before the launch.  It is raised to the device's limit, so the setting is
// the same whichever simulation asks first.
static inline void allow_lds(int lds_max, const void* f, size_t bytes) {
    if (bytes > 65536) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
}

}  // namespace ss

using namespace ss;

struct ss_sim {
    ss_params p;
    int prec = SS_FP64;
    int lds_max = 65536;  // dynamic LDS per workgroup the device allows (gfx950: 160 KiB)
    bool sorted = false;
    // plain path with the particles kept in tile order (SS_BIN_PLAIN_TILED):
    // per-tile LDS counts stored with plain stores, drifted particles through
    // a list, a re-sort in the binning of every `resort`-th step
    bool psort = false;
    int psort_lds = 0;
    uint32_t *strays = nullptr, *scnt = nullptr;  // drifted particles' keys per tile stretch, and their number
    // spectral stage as three fused kernels (tiled path, radices 2-5; with
    // the magnetic solve only for a symmetric stress and unpaired); spec_in
    // then holds moments transformed along z and y only
    bool fused = false;
    bool spec_partial = false;
    // tiled path, symmetric stress of 6 components, not paired: sigma_zz is
    // not binned; the solve takes it from the trace (flow_point)
    bool zz_trace = false;
    // paired fused stage: two real fields per complex transform; spec_in then
    // holds the moments packed in pairs (spec_packed)
    bool paired = false;
    bool spec_packed = false;
    uint2* pairs = nullptr;  // (pencil, mirror pencil) with pencil <= mirror
    int npairs = 0;
    int pencils = 8;  // (y, z) pencils per workgroup of the pencil kernel
    int spec_p = 0;   // 16, 32, 48 or 64 when all three axes have that length (stages fixed at compile time)
    bool mag = false;  // magnetic solve runs
    bool use_e = false, use_rho = false, use_b = false, use_gb = false;
    int nsig = 6;
    int nm = 6;     // moment components
    int nout = 6;   // hot inverse components
    int rs = REC_SMALL;
    uint64_t n = 0;
    int nx, ny, nz, np, nt, ns, na;
    size_t esz = 8;
    uint64_t timestep = 0;
    bool initialised = false;
    bool have_noise = false;
    ss::RefNoise* refn = nullptr;  // SS_NOISE_REFERENCE generators and decoder buffers
    double dr = 0.0;               // sqrt(2 D_r dt), the rotate-angle scale (mod.rs:266-269)
    bool have_fields = false;
    bool have_hist = false;
    hipStream_t stream = nullptr;

    // particle state, A and B (B only on the sorted path)
    void *x = nullptr, *y = nullptr, *z = nullptr, *phi = nullptr, *theta = nullptr;
    void *x2 = nullptr, *y2 = nullptr, *z2 = nullptr, *phi2 = nullptr, *theta2 = nullptr;
    lo_t *lx = nullptr, *ly = nullptr, *lz = nullptr, *lx2 = nullptr, *ly2 = nullptr, *lz2 = nullptr;
    uint32_t *id = nullptr, *id2 = nullptr;
    // SS_NOISE_REFERENCE on the tiled path: slot of each particle id in the
    // current state, so that the noise records follow the storage order
    uint32_t* inv = nullptr;
    bool last_resort = false;  // the last step's update changed the storage order
    bool in_ensemble = false;  // borrowed member of an ensemble: stepped only by it
    bool no_stage = false;     // tiled update: cell records read from global memory, not staged in LDS
    // tiled path
    TileGeo tg;
    int resort = 16;             // steps between tile sorts
    uint64_t steps_since_sort = 0;
    // a re-sort is due: the next update writes the particles to state B in
    // the order of offsets toff2 (computed from the last binning's counts)
    bool pending = false;
    uint32_t *tcnt = nullptr, *toff = nullptr, *toff2 = nullptr, *cursor = nullptr, *part = nullptr;
    // tiled path: direction counts of the last pass, the re-sort's block
    // offsets [ntiles][NDC], far arrivals and far block starts [ntiles]
    uint32_t *dcnt = nullptr, *boff = nullptr, *farcnt = nullptr, *farbase = nullptr;
    int nparts = 0;
    long long *mom_main = nullptr, *mom_inflow = nullptr, *kfix = nullptr;
    double fixed_scale = 1.0;    // value of one fixed-point unit
    uint32_t* hist = nullptr;       // counts of the last step's binning
    uint32_t* hist_next = nullptr;  // tiled path: counts of the next step
    bool swap_hist = false;
    // device step counter and captured graphs of one re-sort period
    uint64_t* d_step = nullptr;
    uint32_t step_off = 0;        // offset of the step being enqueued within a capture
    bool capturing = false;
    bool use_graph = true;
    void* x_buf0 = nullptr;       // the first x buffer: tells which state buffer is current
    uint32_t* toff_buf0 = nullptr;  // the first offsets buffer (current with x_buf0)
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};  // [pending][parity]
    // tiled path: captured runs of steps inside one re-sort period (phase,
    // length, pending/parity key), most recently used last
    struct Segment {
        int phase, len, key;
        hipGraphExec_t exec;
    };
    std::vector<Segment> segments;
    std::vector<Segment> seen;  // shapes run once (eagerly); captured when asked for again
    int graph_steps = 0;
    void* kern = nullptr;
    double* diag_table = nullptr;       // [12][na] FP64: full stress kernel (9) + bin-centre n (3)
    unsigned long long* diag_acc = nullptr;
    void* spec_in = nullptr;
    void* spec_out = nullptr;
    void* spec_full = nullptr;
    void* rho = nullptr;
    void* rec = nullptr;
    void *kx = nullptr, *ky = nullptr, *kz = nullptr, *ik2 = nullptr, *kinv = nullptr;
    void* noise = nullptr;
    double* staging = nullptr;  // n x 5 doubles
    double* staging2 = nullptr; // n x 3 doubles
    Fft3d fft;
    std::vector<void*> owned;
};

namespace ss {

template <typename T>
static Geo<T> geo(const ss_sim* s) {
    Geo<T> g;
    g.nx = s->nx;
    g.ny = s->ny;
    g.nz = s->nz;
    g.np = s->np;
    g.nt = s->nt;
    g.ns = s->ns;
    g.na = s->na;
    GW gw = grid_width(s->p);
    g.gx = (T)gw.x;
    g.gy = (T)gw.y;
    g.gz = (T)gw.z;
    g.gphi = (T)gw.phi;
    g.gtheta = (T)gw.theta;
    g.box[0] = (T)s->p.box_size.x;
    g.box[1] = (T)s->p.box_size.y;
    g.box[2] = (T)s->p.box_size.z;
    const T w[5] = {g.gx, g.gy, g.gz, g.gphi, g.gtheta};
    const int n[5] = {g.nx, g.ny, g.nz, g.np, g.nt};
    const double u = sizeof(T) == 4 ? 0x1p-24 : 0x1p-53;
    for (int a = 0; a < 5; ++a) {
        g.rg[a] = T(1) / w[a];  // correctly rounded in T on the host
        const T tol = (T)(4.0 * u * ((double)n[a] + 1.0));
        g.tol[a] = T(0.5) - tol;  // rounded in T, as grid_index_r's test needs
    }
    return g;
}

template <typename T>
static State<T> state_a(const ss_sim* s) {
    return {(T*)s->x, (T*)s->y, (T*)s->z, (T*)s->phi, (T*)s->theta, s->lx, s->ly, s->lz, s->id};
}
template <typename T>
static State<T> state_b(const ss_sim* s) {
    return {(T*)s->x2, (T*)s->y2, (T*)s->z2, (T*)s->phi2, (T*)s->theta2, s->lx2, s->ly2, s->lz2, s->id2};
}

static void swap_ab(ss_sim* s) {
    std::swap(s->x, s->x2);
    std::swap(s->y, s->y2);
    std::swap(s->z, s->z2);
    std::swap(s->phi, s->phi2);
    std::swap(s->theta, s->theta2);
    std::swap(s->lx, s->lx2);
    std::swap(s->ly, s->ly2);
    std::swap(s->lz, s->lz2);
    std::swap(s->id, s->id2);
}

static void swap_sorted(ss_sim* s) {
    swap_ab(s);
    std::swap(s->toff, s->toff2);
    s->pending = false;
}

// inv[id[slot]] = slot for the current state (after every change of order).
static void launch_inverse_perm(ss_sim* s) {
    if (!s->inv) return;
    hipLaunchKernelGGL(k_inverse_perm, dim3(blocks_for(s->n)), dim3(256), 0, s->stream, (const uint32_t*)s->id, s->n,
                       s->inv);
}

template <typename T>
static UpdateParams<T> update_params(const ss_sim* s) {
    const ss_params& p = s->p;
    UpdateParams<T> P;
    P.dt = (T)p.timestep;
    P.drag = (T)p.magnetic_drag;
    P.shape = (T)p.shape;
    P.vol_excl = (T)p.volume_exclusion;
    P.kappa = (T)p.magnetic_reorientation;
    P.default_dir = default_direction(p) ? 1 : 0;
    for (int k = 0; k < 3; ++k) P.dir[k] = (T)p.field_direction[k];
    P.d_trans = (T)p.diffusion_translational;
    P.dr = (T)sqrt(2. * p.diffusion_rotational * p.timestep);
    P.box[0] = (T)p.box_size.x;
    P.box[1] = (T)p.box_size.y;
    P.box[2] = (T)p.box_size.z;
    P.use_e = s->use_e;
    P.use_rho = s->use_rho;
    P.use_b = s->use_b;
    P.use_gb = s->use_gb;
    P.quasi2d = (p.flags & SS_FLAG_QUASI2D) ? 1 : 0;
    return P;
}

static SolveFlags solve_flags(const ss_sim* s, bool full) {
    SolveFlags f;
    f.nsig = s->nsig;
    f.mag = s->mag ? 1 : 0;
    f.full = full ? 1 : 0;
    f.want_e = s->use_e;
    f.want_b = s->use_b;
    f.want_gb = s->use_gb;
    f.zz_trace = s->zz_trace ? 1 : 0;
    return f;
}

// Scale factors: f = count / (gx gy gz gphi gtheta N) * (Lx Ly Lz)
// (distribution/mod.rs:183, src/bin/simulation/mod.rs:255-257), then
// measure = gphi gtheta (stress.rs:83) or dthph = gtheta gphi (mod.rs:315).
static void scales(const ss_sim* s, double& moment, double& density) {
    GW gw = grid_width(s->p);
    double den = gw.x * gw.y * gw.z * gw.phi * gw.theta * (double)s->n;
    double vol = s->p.box_size.x * s->p.box_size.y * s->p.box_size.z;
    moment = vol / den * (gw.phi * gw.theta);
    density = vol / den * (gw.theta * gw.phi);
}

static void fill_u32(ss_sim* s, uint32_t* p, size_t n, uint32_t v) {
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_u32, dim3(std::max(1u, grid)), dim3(256), 0, s->stream, p, n, v);
}

// Moments and density from the count grid (plain path).
template <typename T>
static void contract_hist(ss_sim* s, const Geo<T>& g, double ms, double ds, T* rho) {
    const size_t per_wg = (size_t)CONTRACT_GROUPS * 256 / CONTRACT_LANES;  // cells per workgroup
    const unsigned nb = (unsigned)((s->ns + per_wg - 1) / per_wg);
    const size_t lds = (size_t)s->nm * s->na * sizeof(T);
    auto go = [&](auto kern) {
        allow_lds(s->lds_max, (const void*)kern, lds);
        hipLaunchKernelGGL(kern, dim3(nb), dim3(256), lds, s->stream, s->hist, g, (const T*)s->kern, (T)ms, (T)ds,
                           (cplx<T>*)s->spec_in, rho);
    };
    if (s->nm == 6)
        go(k_contract_hist<T, 6>);
    else if (s->nm == 9)
        go(k_contract_hist<T, 9>);
    else
        go(k_contract_hist<T, 12>);
}

template <typename T>
static hipError_t launch_bin(ss_sim* s) {
    Geo<T> g = geo<T>(s);
    double ms, ds;
    scales(s, ms, ds);
    T* rho = s->use_rho ? (T*)s->rho : nullptr;
    if (s->psort) {
        // a re-sort is due: count the particles per tile while binning, then
        // move them to the new tile order before the update
        const bool resort = s->steps_since_sort >= (uint64_t)s->resort;
        State<T> st = state_a<T>(s);
        const dim3 grid((unsigned)s->tg.ntiles);
        const size_t lds = (size_t)s->psort_lds;
        auto bin = [&](auto cc) {
            constexpr bool C = decltype(cc)::value;
            allow_lds(s->lds_max, (const void*)k_bin_tile<T, C>, lds);
            hipLaunchKernelGGL((k_bin_tile<T, C>), grid, dim3(256), lds, s->stream, st, (const uint32_t*)s->toff, g,
                               s->tg, s->hist, s->strays, s->scnt, s->tcnt);
            hipLaunchKernelGGL((k_bin_strays<C>), grid, dim3(64), 0, s->stream, (const uint32_t*)s->strays,
                               (const uint32_t*)s->toff, (const uint32_t*)s->scnt, s->ny, s->nz, s->na, s->tg,
                               s->hist, s->tcnt);
        };
        if (!resort) {
            bin(std::false_type());
        } else {
            bin(std::true_type());
            const int nt = s->tg.ntiles;
            hipLaunchKernelGGL(k_scan_reduce, dim3(s->nparts), dim3(256), 0, s->stream, s->tcnt, nt, s->part);
            hipLaunchKernelGGL(k_scan_apply, dim3(s->nparts), dim3(256), 0, s->stream, s->tcnt, nt, s->part,
                               s->nparts, s->toff2, s->cursor);
            hipLaunchKernelGGL(k_psort_permute<T>, grid, dim3(256), 0, s->stream, st, state_b<T>(s),
                               (const uint32_t*)s->toff, (const uint32_t*)s->toff2, g, s->tg, s->cursor);
            swap_sorted(s);
            launch_inverse_perm(s);
            s->steps_since_sort = 0;
        }
        contract_hist<T>(s, g, ms, ds, rho);
    } else if (!s->sorted) {
        fill_u32(s, s->hist, (size_t)s->ns * s->na, 0u);
        hipLaunchKernelGGL(k_bin_hist<T>, dim3(blocks_for(s->n)), dim3(256), 0, s->stream, (const T*)s->x,
                           (const T*)s->y, (const T*)s->z, (const T*)s->phi, (const T*)s->theta, s->n, g,
                           s->hist);
        contract_hist<T>(s, g, ms, ds, rho);
    } else {
        // the moments were accumulated by the previous update (or the initial
        // binning); convert them
        const double sc = s->fixed_scale * ms;
        const double rs = ds;
        const dim3 grid(blocks_for(s->ns));
        if (s->nm == 6)
            hipLaunchKernelGGL((k_tile_convert<T, 6>), grid, dim3(256), 0, s->stream, s->mom_main, s->mom_inflow,
                               s->ns, sc, rs, (cplx<T>*)s->spec_in, rho, s->zz_trace ? 1 : 0);
        else if (s->nm == 9)
            hipLaunchKernelGGL((k_tile_convert<T, 9>), grid, dim3(256), 0, s->stream, s->mom_main, s->mom_inflow,
                               s->ns, sc, rs, (cplx<T>*)s->spec_in, rho, 0);
        else
            hipLaunchKernelGGL((k_tile_convert<T, 12>), grid, dim3(256), 0, s->stream, s->mom_main,
                               s->mom_inflow, s->ns, sc, rs, (cplx<T>*)s->spec_in, rho, 0);
    }
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_solve(ss_sim* s, bool full, void* out) {
    Geo<T> g = geo<T>(s);
    hipLaunchKernelGGL(k_solve<T>, dim3(blocks_for(s->ns)), dim3(256), 0, s->stream,
                       (const cplx<T>*)s->spec_in, (cplx<T>*)out, g, (const T*)s->kx, (const T*)s->ky,
                       (const T*)s->kz, (const T*)s->ik2, (const T*)s->kinv, (T)s->p.hydro_screening,
                       (T)(double)s->ns, solve_flags(s, full));
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_pack(ss_sim* s) {
    hipLaunchKernelGGL(k_pack<T>, dim3(blocks_for(s->ns)), dim3(256), 0, s->stream,
                       (const cplx<T>*)s->spec_out, s->use_rho ? (const T*)s->rho : nullptr, s->ns, s->rs,
                       solve_flags(s, false), (T)s->p.magnetic_dipole_dipole, (T*)s->rec);
    return hipGetLastError();
}

static SpecAxis spec_axis(const FftAxis& a) {
    SpecAxis r;
    r.n = a.n;
    r.nf = a.nf;
    for (int k = 0; k < 16; ++k) r.fac[k] = a.fac[k];
    return r;
}

// Calls f with std::integral_constant<int, P> for the compile-time axis length
// of the fused spectral kernels (0: lengths at run time).
template <typename F>
static auto with_spec_p(int p, F f) {
    switch (p) {
        case 16: return f(std::integral_constant<int, 16>{});
        case 32: return f(std::integral_constant<int, 32>{});
        case 48: return f(std::integral_constant<int, 48>{});
        case 64: return f(std::integral_constant<int, 64>{});
        default: return f(std::integral_constant<int, 0>{});
    }
}

template <typename T>
static hipError_t launch_spec_planes_fwd(ss_sim* s) {
    double ms, ds;
    scales(s, ms, ds);
    const size_t lds = spec_plane_lds(s->ny, s->nz, sizeof(T), s->spec_p > 0);
    // + the count component when the density is wanted (otherwise the update
    // never adds to its inflow, and there is nothing to convert or clear)
    const dim3 grid((unsigned)(s->nx * (s->nm + (s->use_rho ? 1 : 0))));
    T* rho = s->use_rho ? (T*)s->rho : nullptr;
    const SpecAxis ay = spec_axis(s->fft.ax[1]), az = spec_axis(s->fft.ax[2]);
    const cplx<T>* twy = (const cplx<T>*)s->fft.ax[1].tw;
    const cplx<T>* twz = (const cplx<T>*)s->fft.ax[2].tw;
    return with_spec_p(s->spec_p, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        if (s->nm == 6)
            hipLaunchKernelGGL((k_spec_planes_fwd<T, 6, P>), grid, dim3(SPEC_PLANE_THREADS), lds, s->stream,
                               s->mom_main, s->mom_inflow, s->nx, s->fixed_scale * ms, ds, ay, az, twy, twz,
                               (cplx<T>*)s->spec_in, rho, (T*)s->rec, s->rs, s->zz_trace ? 1 : 0);
        else
            hipLaunchKernelGGL((k_spec_planes_fwd<T, 9, P>), grid, dim3(SPEC_PLANE_THREADS), lds, s->stream,
                               s->mom_main, s->mom_inflow, s->nx, s->fixed_scale * ms, ds, ay, az, twy, twz,
                               (cplx<T>*)s->spec_in, rho, (T*)s->rec, s->rs, s->zz_trace ? 1 : 0);
        return hipGetLastError();
    });
}

// The fields after u and W that the fused stage produces (SPEC_OUT_*).
static int spec_outs(const ss_sim* s) {
    return (s->use_e ? SPEC_OUT_E : 0) | (s->mag && s->use_b ? SPEC_OUT_B : 0) | (s->mag && s->use_gb ? SPEC_OUT_GB : 0);
}

template <typename T>
static hipError_t launch_spec_pencils(ss_sim* s) {
    const int np = s->ny * s->nz;
    const int nout = s->nout, nmag = s->mag ? 3 : 0;
    const size_t lds = spec_pencil_lds(s->nx, std::max(s->nsig + nmag, nout) * s->pencils, sizeof(T));
    const dim3 grid((unsigned)((np + s->pencils - 1) / s->pencils));
    return with_spec_p(s->spec_p, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        hipLaunchKernelGGL((k_spec_pencils<T, P>), grid, dim3(SPEC_PENCIL_THREADS), lds, s->stream,
                           (const cplx<T>*)s->spec_in, (cplx<T>*)s->spec_out, np, s->nz, s->pencils, s->nsig, nout,
                           spec_axis(s->fft.ax[0]), (const cplx<T>*)s->fft.ax[0].tw, (const T*)s->kx,
                           (const T*)s->ky, (const T*)s->kz, (const T*)s->ik2, (const T*)s->kinv,
                           (T)s->p.hydro_screening, (T)(double)s->ns, s->zz_trace ? 1 : 0, nmag, spec_outs(s));
        return hipGetLastError();
    });
}

template <typename T>
static hipError_t launch_spec_planes_fwd2(ss_sim* s) {
    double ms, ds;
    scales(s, ms, ds);
    const size_t lds = spec_plane_lds(s->ny, s->nz, sizeof(T), s->spec_p > 0);
    const int npk = (s->nm + 1) / 2;
    const dim3 grid((unsigned)(s->nx * (npk + 1)));  // + the count component
    T* rho = s->use_rho ? (T*)s->rho : nullptr;
    const SpecAxis ay = spec_axis(s->fft.ax[1]), az = spec_axis(s->fft.ax[2]);
    const cplx<T>* twy = (const cplx<T>*)s->fft.ax[1].tw;
    const cplx<T>* twz = (const cplx<T>*)s->fft.ax[2].tw;
    return with_spec_p(s->spec_p, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        if (s->nm == 6)
            hipLaunchKernelGGL((k_spec_planes_fwd2<T, 6, P>), grid, dim3(SPEC_PLANE_THREADS), lds, s->stream,
                               s->mom_main, s->mom_inflow, s->nx, s->fixed_scale * ms, ds, ay, az, twy, twz,
                               (cplx<T>*)s->spec_in, rho, (T*)s->rec, s->rs);
        else
            hipLaunchKernelGGL((k_spec_planes_fwd2<T, 9, P>), grid, dim3(SPEC_PLANE_THREADS), lds, s->stream,
                               s->mom_main, s->mom_inflow, s->nx, s->fixed_scale * ms, ds, ay, az, twy, twz,
                               (cplx<T>*)s->spec_in, rho, (T*)s->rec, s->rs);
        return hipGetLastError();
    });
}

template <typename T>
static hipError_t launch_spec_pencils2(ss_sim* s) {
    const int np = s->ny * s->nz;
    const int nout = s->use_e ? 12 : 6;
    const int npk = (s->nsig + 1) / 2;
    const size_t lds = spec_pencil_lds(s->nx, 2 * std::max(npk, nout / 2) * s->pencils, sizeof(T));
    const dim3 grid((unsigned)((s->npairs + s->pencils - 1) / s->pencils));
    return with_spec_p(s->spec_p, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        hipLaunchKernelGGL((k_spec_pencils2<T, P>), grid, dim3(SPEC_PENCIL_THREADS), lds, s->stream,
                           (const cplx<T>*)s->spec_in, (cplx<T>*)s->spec_out, (const uint2*)s->pairs, s->npairs, np,
                           s->nz, s->ny, s->pencils, s->nsig, nout, spec_axis(s->fft.ax[0]),
                           (const cplx<T>*)s->fft.ax[0].tw, (const T*)s->kx, (const T*)s->ky, (const T*)s->kz,
                           (const T*)s->ik2, (const T*)s->kinv, (T)s->p.hydro_screening, (T)(double)s->ns);
        return hipGetLastError();
    });
}

template <typename T>
static hipError_t launch_spec_planes_inv2(ss_sim* s) {
    const int nout = s->use_e ? 12 : 6;
    const size_t lds = spec_plane_lds(s->ny, s->nz, sizeof(T), s->spec_p > 0);
    const dim3 grid((unsigned)(s->nx * (nout / 2)));
    return with_spec_p(s->spec_p, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        hipLaunchKernelGGL((k_spec_planes_inv2<T, P>), grid, dim3(SPEC_PLANE_THREADS), lds, s->stream,
                           (const cplx<T>*)s->spec_out, s->nx, spec_axis(s->fft.ax[1]), spec_axis(s->fft.ax[2]),
                           (const cplx<T>*)s->fft.ax[1].tw, (const cplx<T>*)s->fft.ax[2].tw, (T*)s->rec, s->rs);
        return hipGetLastError();
    });
}

template <typename T>
static hipError_t launch_spec_planes_inv(ss_sim* s) {
    const int nout = s->nout;
    const size_t lds = spec_plane_lds(s->ny, s->nz, sizeof(T), s->spec_p > 0);
    const dim3 grid((unsigned)(s->nx * nout));
    return with_spec_p(s->spec_p, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        hipLaunchKernelGGL((k_spec_planes_inv<T, P>), grid, dim3(SPEC_PLANE_THREADS), lds, s->stream,
                           (const cplx<T>*)s->spec_out, s->nx, spec_axis(s->fft.ax[1]), spec_axis(s->fft.ax[2]),
                           (const cplx<T>*)s->fft.ax[1].tw, (const cplx<T>*)s->fft.ax[2].tw, (T*)s->rec, s->rs,
                           spec_outs(s), (T)s->p.magnetic_dipole_dipole);
        return hipGetLastError();
    });
}

// Bytes per window slot of the tiled update's staged cell records, or 0 when
// the records are read from global memory (full 32-value records, a window
// that leaves no LDS for them, or switched off for an ensemble).
static size_t tile_rec_bytes(const ss_sim* s) {
    if (s->rs != REC_SMALL || s->no_stage) return 0;
    const size_t b = (size_t)REC_SMALL * (s->prec == SS_FP32 ? sizeof(float) : sizeof(double));
    return tile_lds_bytes(s->tg.ws, s->nm, s->na, s->nx + s->ny + s->nz, b) <= (size_t)s->lds_max ? b : 0;
}

template <typename T, int NM, bool MOVE, int NZ, bool KEEP>
static void launch_tile_kernel(ss_sim* s, const TileArgs<T, NM>& a) {
    const int per = (s->tg.ntiles + s->tg.nxcd - 1) / s->tg.nxcd;
    const dim3 grid((unsigned)(per * s->tg.nxcd));
    const size_t lds = tile_lds_bytes(s->tg.ws, NM, s->na, s->nx + s->ny + s->nz, a.stage ? tile_rec_bytes(s) : 0);
    auto go = [&](auto kern) {
        allow_lds(s->lds_max, (const void*)kern, lds);
        hipLaunchKernelGGL(kern, grid, dim3(256), lds, s->stream, a);
    };
    // the default physics (no magnetic, strain, density or quasi-2D terms)
    // with those terms compiled out
    if constexpr (MOVE && NZ == NZ_PHILOX && NM == 6) {
        if (!full_physics(a.P)) {
            if (a.dtoff)
                go(k_tile_update<T, NM, MOVE, NZ, KEEP, true, false>);
            else
                go(k_tile_update<T, NM, MOVE, NZ, KEEP, false, false>);
            return;
        }
    }
    if (MOVE && a.dtoff)
        go(k_tile_update<T, NM, MOVE, NZ, KEEP, MOVE>);
    else
        go(k_tile_update<T, NM, MOVE, NZ, KEEP, false>);
}

template <typename T, int NM, bool MOVE>
static void launch_tile_nm(ss_sim* s, const TileArgs<T, NM>& a) {
    const bool keep = a.hist != nullptr;
    if (s->p.noise_mode == SS_NOISE_EXTERNAL) {
        if (keep) launch_tile_kernel<T, NM, MOVE, NZ_EXT, true>(s, a);
        else launch_tile_kernel<T, NM, MOVE, NZ_EXT, false>(s, a);
    } else if (s->p.noise_mode == SS_NOISE_REFERENCE) {
        if (keep) launch_tile_kernel<T, NM, MOVE, NZ_REC, true>(s, a);
        else launch_tile_kernel<T, NM, MOVE, NZ_REC, false>(s, a);
    } else {
        if (keep) launch_tile_kernel<T, NM, MOVE, NZ_PHILOX, true>(s, a);
        else launch_tile_kernel<T, NM, MOVE, NZ_PHILOX, false>(s, a);
    }
}

template <typename T, int NM>
static TileArgs<T, NM> tile_args(ss_sim* s) {
    TileArgs<T, NM> a;
    a.s = state_a<T>(s);
    a.toff = s->toff;
    a.n = s->n;
    a.g = geo<T>(s);
    a.tg = s->tg;
    a.P = update_params<T>(s);
    a.rec = (const T*)s->rec;
    a.rs = s->rs;
    a.seed = s->p.seed;
    a.step = s->d_step;
    a.step_off = s->step_off;
    a.noise = (const T*)s->noise;
    a.kfix = s->kfix;
    a.main = s->mom_main;
    a.inflow = s->mom_inflow;
    a.hist = (s->p.flags & SS_FLAG_KEEP_DISTRIBUTION) ? s->hist_next : nullptr;
    a.d = state_b<T>(s);
    a.dtoff = s->pending ? s->toff2 : nullptr;
    a.dcnt = s->dcnt;
    a.boff = s->boff;
    a.farbase = s->farbase;
    a.farctr = s->cursor;
    a.stage = tile_rec_bytes(s) > 0 ? 1 : 0;
    a.zz_trace = s->zz_trace ? 1 : 0;
    return a;
}

// Tiled path: move every particle and bin the new state (MOVE), or only bin
// the current state (initial sampling).  A pending re-sort makes the move
// write state B in the new order, which then becomes current.
template <typename T, bool MOVE>
static hipError_t launch_tile_update(ss_sim* s) {
    if (s->hist_next) fill_u32(s, s->hist_next, (size_t)s->ns * s->na, 0u);
    if (!MOVE && s->pending) return hipErrorInvalidValue;
    if (s->nm == 6)
        launch_tile_nm<T, 6, MOVE>(s, tile_args<T, 6>(s));
    else if (s->nm == 9)
        launch_tile_nm<T, 9, MOVE>(s, tile_args<T, 9>(s));
    else
        launch_tile_nm<T, 12, MOVE>(s, tile_args<T, 12>(s));
    if (MOVE) s->last_resort = s->pending;
    if (MOVE && s->pending) {
        swap_sorted(s);
        launch_inverse_perm(s);
    }
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_update(ss_sim* s) {
    if (s->sorted) return launch_tile_update<T, true>(s);
    Geo<T> g = geo<T>(s);
    UpdateParams<T> P = update_params<T>(s);
    State<T> st = state_a<T>(s);
    dim3 grid(blocks_for(s->n));
    if (s->p.noise_mode == SS_NOISE_EXTERNAL)
        hipLaunchKernelGGL((k_update_plain<T, NZ_EXT>), grid, dim3(256), 0, s->stream, st, s->n, g, P,
                           (const T*)s->rec, s->rs, s->p.seed, s->d_step, s->step_off, (const T*)s->noise);
    else if (s->p.noise_mode == SS_NOISE_REFERENCE)
        hipLaunchKernelGGL((k_update_plain<T, NZ_REC>), grid, dim3(256), 0, s->stream, st, s->n, g, P,
                           (const T*)s->rec, s->rs, s->p.seed, s->d_step, s->step_off, (const T*)s->noise);
    else if (full_physics(P))
        hipLaunchKernelGGL((k_update_plain<T, NZ_PHILOX>), grid, dim3(256), 0, s->stream, st, s->n, g, P,
                           (const T*)s->rec, s->rs, s->p.seed, s->d_step, s->step_off, (const T*)s->noise);
    else  // the default physics with the optional terms compiled out
        hipLaunchKernelGGL((k_update_plain<T, NZ_PHILOX, false>), grid, dim3(256), 0, s->stream, st, s->n, g, P,
                           (const T*)s->rec, s->rs, s->p.seed, s->d_step, s->step_off, (const T*)s->noise);
    return hipGetLastError();
}

// Offsets of the next re-sort from the direction counts of the last pass
// (k_dir_totals, the tiles' exclusive scan into toff2, k_dir_offsets; far
// arrivals counted by k_far_count first).  The next update places the
// particles (pending).
static hipError_t launch_count_scan(ss_sim* s) {
    const int nt = s->tg.ntiles;
    const dim3 g1((unsigned)((nt + 3) / 4));  // a wavefront per tile
    fill_u32(s, s->farcnt, (size_t)nt, 0u);
    if (s->prec == SS_FP32)
        hipLaunchKernelGGL(k_far_count<float>, dim3(nt), dim3(256), 0, s->stream, state_a<float>(s),
                           (const uint32_t*)s->toff, geo<float>(s), s->tg, (const uint32_t*)s->dcnt, s->farcnt);
    else
        hipLaunchKernelGGL(k_far_count<double>, dim3(nt), dim3(256), 0, s->stream, state_a<double>(s),
                           (const uint32_t*)s->toff, geo<double>(s), s->tg, (const uint32_t*)s->dcnt, s->farcnt);
    hipLaunchKernelGGL(k_dir_totals, g1, dim3(256), 0, s->stream, (const uint32_t*)s->dcnt,
                       (const uint32_t*)s->farcnt, s->tg, s->tcnt);
    hipLaunchKernelGGL(k_scan_reduce, dim3(s->nparts), dim3(256), 0, s->stream, s->tcnt, nt, s->part);
    hipLaunchKernelGGL(k_scan_apply, dim3(s->nparts), dim3(256), 0, s->stream, s->tcnt, nt, s->part, s->nparts,
                       s->toff2, s->cursor);
    hipLaunchKernelGGL(k_dir_offsets, g1, dim3(256), 0, s->stream, (const uint32_t*)s->dcnt,
                       (const uint32_t*)s->toff2, s->tg, s->boff, s->farbase, s->cursor);
    s->pending = true;
    s->steps_since_sort = 0;
    return hipGetLastError();
}

// Counting sort of particles in arbitrary order by tile (ss_init, resume).
template <typename T>
static hipError_t launch_initial_sort(ss_sim* s) {
    const int nt = s->tg.ntiles;
    fill_u32(s, s->tcnt, (size_t)nt, 0u);
    hipLaunchKernelGGL(k_tile_count_state<T>, dim3(blocks_for(s->n)), dim3(256), 0, s->stream, state_a<T>(s),
                       s->n, geo<T>(s), s->tg, s->tcnt);
    // the offsets buffer that becomes current pairs with the state buffer that
    // becomes current, as every later re-sort swaps both
    const bool to_b0 = s->x != s->x_buf0;
    if ((s->toff == s->toff_buf0) != to_b0) std::swap(s->toff, s->toff2);
    hipLaunchKernelGGL(k_scan_reduce, dim3(s->nparts), dim3(256), 0, s->stream, s->tcnt, nt, s->part);
    hipLaunchKernelGGL(k_scan_apply, dim3(s->nparts), dim3(256), 0, s->stream, s->tcnt, nt, s->part, s->nparts,
                       s->toff, s->cursor);
    hipLaunchKernelGGL(k_tile_permute<T>, dim3(blocks_for(s->n)), dim3(256), 0, s->stream, state_a<T>(s),
                       state_b<T>(s), s->n, geo<T>(s), s->tg, s->toff, s->cursor);
    hipError_t e = hipGetLastError();
    swap_ab(s);
    launch_inverse_perm(s);
    s->steps_since_sort = 0;
    s->pending = false;
    return e;
}

enum Stage { ST_BIN = 0, ST_FWD, ST_SOLVE, ST_INV, ST_PACK, ST_NOISE, ST_UPDATE, ST_SORT, ST_COUNT };
static const char* STAGE_NAMES[ST_COUNT] = {"bin", "fft_forward", "solve", "fft_inverse",
                                            "pack", "noise", "update", "sort"};

// Moments -> cell records (bin/convert, forward FFT, solve, inverse FFT,
// pack).  ev (may be null): events recorded after each stage.
template <typename T>
static int spectral_stage(ss_sim* s, hipEvent_t* ev) {
    hipError_t e;
    auto mark = [&](int st) {
        if (ev) (void)hipEventRecord(ev[st + 1], s->stream);
    };
    if (s->fused) {
        // stages: bin = moments -> z, y transforms; solve = x transforms and
        // solve; fft_inverse = y, z inverse transforms into the records
        if ((e = s->paired ? launch_spec_planes_fwd2<T>(s) : launch_spec_planes_fwd<T>(s)) != hipSuccess)
            return fail(SS_ERR_DEVICE, std::string("spectral planes: ") + hipGetErrorString(e));
        mark(ST_BIN);
        mark(ST_FWD);
        if ((e = s->paired ? launch_spec_pencils2<T>(s) : launch_spec_pencils<T>(s)) != hipSuccess)
            return fail(SS_ERR_DEVICE, std::string("spectral pencils: ") + hipGetErrorString(e));
        mark(ST_SOLVE);
        if ((e = s->paired ? launch_spec_planes_inv2<T>(s) : launch_spec_planes_inv<T>(s)) != hipSuccess)
            return fail(SS_ERR_DEVICE, std::string("spectral planes: ") + hipGetErrorString(e));
        mark(ST_INV);
        mark(ST_PACK);
        s->spec_partial = true;
        s->spec_packed = s->paired;
    } else {
        if ((e = launch_bin<T>(s)) != hipSuccess)
            return fail(SS_ERR_DEVICE, std::string("bin: ") + hipGetErrorString(e));
        mark(ST_BIN);
        int rc = s->fft.run(s->spec_in, s->nm, -1, s->stream);
        if (rc != SS_OK) return rc;
        mark(ST_FWD);
        if ((e = launch_solve<T>(s, false, s->spec_out)) != hipSuccess)
            return fail(SS_ERR_DEVICE, std::string("solve: ") + hipGetErrorString(e));
        mark(ST_SOLVE);
        rc = s->fft.run(s->spec_out, s->nout, +1, s->stream);
        if (rc != SS_OK) return rc;
        mark(ST_INV);
        if ((e = launch_pack<T>(s)) != hipSuccess)
            return fail(SS_ERR_DEVICE, std::string("pack: ") + hipGetErrorString(e));
        mark(ST_PACK);
    }
    return SS_OK;
}

// One do_timestep.  ev (may be null): events recorded after each stage.
template <typename T>
static int step_once(ss_sim* s, hipEvent_t* ev) {
    hipError_t e;
    auto mark = [&](int st) {
        if (ev) (void)hipEventRecord(ev[st + 1], s->stream);
    };
    if (ev) (void)hipEventRecord(ev[0], s->stream);
    if (s->swap_hist) std::swap(s->hist, s->hist_next);
    int rc = spectral_stage<T>(s, ev);
    if (rc != SS_OK) return rc;
    if (s->refn) {
        // the RandomVectors of do_timestep (mod.rs:259-286) from the reference's streams
        if ((e = refnoise_enqueue(s->refn, s->noise, s->prec == SS_FP64, s->dr, s->inv, s->stream)) != hipSuccess)
            return fail(SS_ERR_DEVICE, std::string("reference noise: ") + hipGetErrorString(e));
        s->have_noise = true;
    }
    mark(ST_NOISE);
    if ((e = launch_update<T>(s)) != hipSuccess)
        return fail(SS_ERR_DEVICE, std::string("update: ") + hipGetErrorString(e));
    mark(ST_UPDATE);
    if (s->sorted) {
        s->steps_since_sort += 1;
        if (s->steps_since_sort >= (uint64_t)s->resort) {
            if ((e = launch_count_scan(s)) != hipSuccess)
                return fail(SS_ERR_DEVICE, std::string("sort: ") + hipGetErrorString(e));
        }
        s->swap_hist = s->hist_next != nullptr;
    } else if (s->psort) {
        s->steps_since_sort += 1;  // the binning of a later step re-sorts
    }
    mark(ST_SORT);
    if (s->capturing) {
        s->step_off += 1;
    } else {
        hipLaunchKernelGGL(k_add_step, dim3(1), dim3(64), 0, s->stream, s->d_step, 1u);
    }
    s->timestep += 1;
    s->have_fields = true;
    s->have_hist = !s->sorted || (s->p.flags & SS_FLAG_KEEP_DISTRIBUTION);
    return SS_OK;
}

// Sample the distribution of the current particles (Simulation::init,
// src/bin/simulation/mod.rs:190-195).  Tiled path: sort by tile, then bin
// without moving; the next step converts the accumulated moments.
template <typename T>
static int initial_binning(ss_sim* s) {
    hipError_t e;
    if (s->sorted) {
        if ((e = launch_initial_sort<T>(s)) != hipSuccess) return fail(SS_ERR_DEVICE, hipGetErrorString(e));
        SS_HIP_CHECK(hipMemsetAsync(s->mom_inflow, 0, (size_t)s->ns * (s->nm + 1) * 8, s->stream));
        if ((e = launch_tile_update<T, false>(s)) != hipSuccess) return fail(SS_ERR_DEVICE, hipGetErrorString(e));
        if (s->hist_next) std::swap(s->hist, s->hist_next);
        s->swap_hist = false;
    } else {
        if (s->psort && (e = launch_initial_sort<T>(s)) != hipSuccess)
            return fail(SS_ERR_DEVICE, hipGetErrorString(e));
        if ((e = launch_bin<T>(s)) != hipSuccess) return fail(SS_ERR_DEVICE, hipGetErrorString(e));
    }
    s->have_hist = !s->sorted || (s->p.flags & SS_FLAG_KEEP_DISTRIBUTION);
    return SS_OK;
}

template <typename T>
static int import_particles(ss_sim* s, const double* parts, const double* comp, bool sanitize) {
    SS_HIP_CHECK(hipMemcpyAsync(s->staging, parts, sizeof(double) * 5 * s->n, hipMemcpyHostToDevice, s->stream));
    if (comp)
        SS_HIP_CHECK(hipMemcpyAsync(s->staging2, comp, sizeof(double) * 3 * s->n, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(k_import<T>, dim3(blocks_for(s->n)), dim3(256), 0, s->stream, s->staging,
                       comp ? s->staging2 : nullptr, s->n, state_a<T>(s), s->p.box_size.x, s->p.box_size.y,
                       s->p.box_size.z, sanitize ? 1 : 0);
    SS_HIP_CHECK(hipGetLastError());
    return initial_binning<T>(s);
}

template <typename T>
static int export_particles(ss_sim* s, double* out, double* comp, uint64_t n) {
    hipLaunchKernelGGL(k_export<T>, dim3(blocks_for(s->n)), dim3(256), 0, s->stream, state_a<T>(s), s->n,
                       s->staging, comp ? s->staging2 : nullptr, s->p.box_size.x, s->p.box_size.y,
                       s->p.box_size.z);
    SS_HIP_CHECK(hipGetLastError());
    if (out) SS_HIP_CHECK(hipMemcpyAsync(out, s->staging, sizeof(double) * 5 * n, hipMemcpyDeviceToHost, s->stream));
    if (comp) SS_HIP_CHECK(hipMemcpyAsync(comp, s->staging2, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s->stream));
    SS_HIP_CHECK(hipStreamSynchronize(s->stream));
    return SS_OK;
}

// Full (non-hot) solve from the retained forward-transformed moments into
// spec_full: u(3), G(9), [b(3), gradB(9)], inverse transformed.
template <typename T>
static int full_fields(ss_sim* s) {
    if (s->spec_packed) {  // the paired stage left the moments packed in pairs
        const size_t ns = (size_t)s->ns;
        hipLaunchKernelGGL(k_spec_unpack<T>, dim3(blocks_for(ns)), dim3(256), 0, s->stream,
                           (const cplx<T>*)s->spec_in, s->nsig, s->nx, s->ny, s->nz, (cplx<T>*)s->spec_full);
        SS_HIP_CHECK(hipGetLastError());
        SS_HIP_CHECK(hipMemcpyAsync(s->spec_in, s->spec_full, sizeof(cplx<T>) * ns * s->nsig, hipMemcpyDeviceToDevice,
                                    s->stream));
        s->spec_packed = false;
    }
    if (s->spec_partial) {  // the fused step left the x transform of the moments undone
        int rc = s->fft.run_axis(s->spec_in, s->nm, -1, 0, s->stream);
        if (rc != SS_OK) return rc;
        s->spec_partial = false;
    }
    hipError_t e = launch_solve<T>(s, true, s->spec_full);
    if (e != hipSuccess) return fail(SS_ERR_DEVICE, hipGetErrorString(e));
    const int nfull = 12 + (s->mag ? 12 : 0);
    return s->fft.run(s->spec_full, nfull, +1, s->stream);
}

template <typename T>
static int copy_real(ss_sim* s, const void* spec, int comp0, int ncomp, double* host) {
    const size_t cnt = (size_t)ncomp * s->ns;
    double* tmp = s->staging;  // n*5 doubles; sized at create to hold the field copies too
    hipLaunchKernelGGL(k_real_parts<T>, dim3(blocks_for(cnt)), dim3(256), 0, s->stream,
                       (const cplx<T>*)spec + (size_t)comp0 * s->ns, cnt, tmp);
    SS_HIP_CHECK(hipGetLastError());
    SS_HIP_CHECK(hipMemcpyAsync(host, tmp, sizeof(double) * cnt, hipMemcpyDeviceToHost, s->stream));
    SS_HIP_CHECK(hipStreamSynchronize(s->stream));
    return SS_OK;
}

// Steps per captured graph: one re-sort period on the tiled path (its first
// update carries out the re-sort prepared at the end of the previous period,
// swapping the state buffers), 8 steps on the plain path.
static int graph_period(const ss_sim* s) { return s->sorted ? s->resort : 8; }

// Capture nsteps steps from the current state into *out.  Capturing
// enqueues nothing; it advances the host-side bookkeeping exactly as running
// the steps would (step_once), so the caller must launch the graph once
// afterwards.
template <typename T>
static int capture_steps(ss_sim* s, int nsteps, hipGraphExec_t* out) {
    SS_HIP_CHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    s->capturing = true;
    s->step_off = 0;
    int rc = SS_OK;
    for (int k = 0; k < nsteps && rc == SS_OK; ++k) rc = step_once<T>(s, nullptr);
    if (rc == SS_OK) hipLaunchKernelGGL(k_add_step, dim3(1), dim3(64), 0, s->stream, s->d_step, (uint32_t)nsteps);
    s->capturing = false;
    s->step_off = 0;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    if (rc != SS_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(SS_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(SS_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    return SS_OK;
}

// One re-sort period (the tiled path) or 8 steps (the plain paths) as the
// graph of period key `key`.
template <typename T>
static int capture_period(ss_sim* s, int key) {
    const int K = graph_period(s);
    const int rc = capture_steps<T>(s, K, &s->graph[key]);
    if (rc == SS_OK) s->graph_steps = K;
    return rc;
}

// The host-side bookkeeping of `len` steps of the tiled path from the start
// of a re-sort period or inside one (what step_once does for them): only a
// period's first update re-sorts (when one is pending), and a run that
// reaches the end of the period computes the next re-sort's offsets.
static void advance_tiled(ss_sim* s, int len) {
    s->timestep += (uint64_t)len;
    s->last_resort = len == 1 && s->pending;
    if (s->pending) swap_sorted(s);
    s->steps_since_sort += (uint64_t)len;
    if (s->steps_since_sort >= (uint64_t)s->resort) {
        s->pending = true;
        s->steps_since_sort = 0;
    }
    s->have_fields = true;
    s->spec_partial = s->fused;
    s->spec_packed = s->fused && s->paired;
}

constexpr size_t MAX_SEGMENTS = 8;

// Destroy a captured graph that may have a launch still queued or running on
// `st`: the executable owns the kernel arguments of its nodes, so it is only
// released once the stream has drained.
static void destroy_exec(hipStream_t st, hipGraphExec_t g) {
    if (!g) return;
    if (st) (void)hipStreamSynchronize(st);
    (void)hipGraphExecDestroy(g);
}

// nsteps steps: whole periods as graph launches; on the tiled path any other
// run of steps inside one re-sort period as a launch of a captured segment
// (captured the second time a run of its phase, length and buffer state is
// asked for; the last MAX_SEGMENTS kept), so that a caller stepping a few
// steps at a time pays one graph launch per call instead of five kernel
// launches per step; the rest eager.
template <typename T>
static int run_steps(ss_sim* s, uint64_t nsteps) {
    const int K = graph_period(s);
    while (nsteps > 0) {
        const bool aligned = !s->sorted || s->steps_since_sort == 0;
        if (s->use_graph && aligned && nsteps >= (uint64_t)K) {
            const int key = (s->pending ? 2 : 0) + (s->x == s->x_buf0 ? 0 : 1);
            if (!s->graph[key]) {
                int rc = capture_period<T>(s, key);
                if (rc != SS_OK) return rc;
            } else {
                // advance the bookkeeping the captured period performs
                for (int k = 0; k < K; ++k) s->timestep += 1;
                if (s->sorted) {
                    s->last_resort = K == 1 && s->pending;  // only a period's first update re-sorts
                    if (s->pending) swap_sorted(s);
                    s->pending = true;
                    s->steps_since_sort = 0;
                }
                s->have_fields = true;
                s->spec_partial = s->fused;
                s->spec_packed = s->fused && s->paired;
            }
            SS_HIP_CHECK(hipGraphLaunch(s->graph[key], s->stream));
            nsteps -= (uint64_t)K;
        } else if (s->use_graph && s->sorted && s->hist_next == nullptr) {
            // the run of steps up to the end of this re-sort period (or of
            // the request) as one captured segment; the first request of a
            // shape runs eagerly (a capture costs more than the launches it
            // saves once), a repeated one is captured and kept
            const int phase = (int)s->steps_since_sort;
            const int len = (int)std::min<uint64_t>(nsteps, (uint64_t)(s->resort - phase));
            const int key = (s->pending ? 2 : 0) + (s->x == s->x_buf0 ? 0 : 1);
            auto same = [&](const ss_sim::Segment& g) { return g.phase == phase && g.len == len && g.key == key; };
            auto it = std::find_if(s->segments.begin(), s->segments.end(), same);
            if (it == s->segments.end() && std::find_if(s->seen.begin(), s->seen.end(), same) == s->seen.end()) {
                if (s->seen.size() >= 4 * MAX_SEGMENTS) s->seen.erase(s->seen.begin());
                s->seen.push_back({phase, len, key, nullptr});
                for (int k = 0; k < len; ++k) {
                    int rc = step_once<T>(s, nullptr);
                    if (rc != SS_OK) return rc;
                }
                nsteps -= (uint64_t)len;
                continue;
            }
            hipGraphExec_t exec = nullptr;
            if (it == s->segments.end()) {
                if (s->segments.size() >= MAX_SEGMENTS) {
                    destroy_exec(s->stream, s->segments.front().exec);
                    s->segments.erase(s->segments.begin());
                }
                int rc = capture_steps<T>(s, len, &exec);
                if (rc != SS_OK) return rc;
                s->segments.push_back({phase, len, key, exec});
            } else {
                exec = it->exec;
                const ss_sim::Segment g = *it;  // most recently used last
                s->segments.erase(it);
                s->segments.push_back(g);
                advance_tiled(s, len);
            }
            SS_HIP_CHECK(hipGraphLaunch(exec, s->stream));
            nsteps -= (uint64_t)len;
        } else {
            int rc = step_once<T>(s, nullptr);
            if (rc != SS_OK) return rc;
            nsteps -= 1;
        }
    }
    return SS_OK;
}

// The host-side state a step of the tiled path changes (step_once, the
// re-sort its update may carry out, the next re-sort's offsets at the end of
// a period), saved and restored around a capture that must not count.
struct StepBook {
    uint64_t timestep, steps_since_sort;
    bool pending, last_resort, have_fields, have_hist, spec_partial, spec_packed, swap_hist;
    void *x, *y, *z, *phi, *theta, *x2, *y2, *z2, *phi2, *theta2;
    lo_t *lx, *ly, *lz, *lx2, *ly2, *lz2;
    uint32_t *id, *id2, *toff, *toff2;
};

static StepBook save_book(const ss_sim* s) {
    return {s->timestep, s->steps_since_sort, s->pending, s->last_resort, s->have_fields, s->have_hist,
            s->spec_partial, s->spec_packed, s->swap_hist, s->x, s->y, s->z, s->phi, s->theta, s->x2, s->y2,
            s->z2, s->phi2, s->theta2, s->lx, s->ly, s->lz, s->lx2, s->ly2, s->lz2, s->id, s->id2, s->toff,
            s->toff2};
}

static void restore_book(ss_sim* s, const StepBook& b) {
    s->timestep = b.timestep;
    s->steps_since_sort = b.steps_since_sort;
    s->pending = b.pending;
    s->last_resort = b.last_resort;
    s->have_fields = b.have_fields;
    s->have_hist = b.have_hist;
    s->spec_partial = b.spec_partial;
    s->spec_packed = b.spec_packed;
    s->swap_hist = b.swap_hist;
    s->x = b.x;
    s->y = b.y;
    s->z = b.z;
    s->phi = b.phi;
    s->theta = b.theta;
    s->x2 = b.x2;
    s->y2 = b.y2;
    s->z2 = b.z2;
    s->phi2 = b.phi2;
    s->theta2 = b.theta2;
    s->lx = b.lx;
    s->ly = b.ly;
    s->lz = b.lz;
    s->lx2 = b.lx2;
    s->ly2 = b.ly2;
    s->lz2 = b.lz2;
    s->id = b.id;
    s->id2 = b.id2;
    s->toff = b.toff;
    s->toff2 = b.toff2;
}

// Capture the graphs that the next run_steps(s, nsteps) replays (whole
// periods and the runs of steps between period boundaries), without running
// anything: the bookkeeping the captures advance is restored afterwards.
// The tiled path with captured graphs only (not with a kept distribution or
// the reference's noise stream, whose host-side generator state a capture
// would advance); elsewhere nothing to do.
template <typename T>
static int prepare_steps(ss_sim* s, uint64_t nsteps) {
    if (!s->use_graph || !s->sorted || s->hist_next != nullptr || s->refn) return SS_OK;
    const StepBook book = save_book(s);
    const int K = graph_period(s);
    std::vector<ss_sim::Segment> made;
    int rc = SS_OK;
    while (nsteps > 0 && rc == SS_OK) {
        const int key = (s->pending ? 2 : 0) + (s->x == s->x_buf0 ? 0 : 1);
        if (s->steps_since_sort == 0 && nsteps >= (uint64_t)K) {
            if (!s->graph[key])
                rc = capture_period<T>(s, key);
            else
                advance_tiled(s, K);
            nsteps -= (uint64_t)K;
            continue;
        }
        const int phase = (int)s->steps_since_sort;
        const int len = (int)std::min<uint64_t>(nsteps, (uint64_t)(s->resort - phase));
        auto same = [&](const ss_sim::Segment& g) { return g.phase == phase && g.len == len && g.key == key; };
        if (std::find_if(s->segments.begin(), s->segments.end(), same) != s->segments.end() ||
            std::find_if(made.begin(), made.end(), same) != made.end()) {
            advance_tiled(s, len);
        } else {
            hipGraphExec_t exec = nullptr;
            rc = capture_steps<T>(s, len, &exec);
            if (rc == SS_OK) made.push_back({phase, len, key, exec});
        }
        nsteps -= (uint64_t)len;
    }
    restore_book(s, book);
    for (auto& g : made) {
        if (rc != SS_OK) {
            (void)hipGraphExecDestroy(g.exec);
            continue;
        }
        if (s->segments.size() >= MAX_SEGMENTS) {
            destroy_exec(s->stream, s->segments.front().exec);
            s->segments.erase(s->segments.begin());
        }
        s->segments.push_back(g);
    }
    return rc;
}

static int alloc(ss_sim* s, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) return fail(SS_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    s->owned.push_back(*p);
    return SS_OK;
}

static void drop_graphs(ss_sim* s) {
    for (int k = 0; k < 4; ++k)
        if (s->graph[k]) {
            destroy_exec(s->stream, s->graph[k]);
            s->graph[k] = nullptr;
        }
    for (auto& g : s->segments) destroy_exec(s->stream, g.exec);
    s->segments.clear();
    s->seen.clear();
}

static void free_all(ss_sim* s) {
    drop_graphs(s);
    for (void* p : s->owned) (void)hipFree(p);
    s->owned.clear();
    s->fft.release();
    refnoise_destroy(s->refn);
    s->refn = nullptr;
    if (s->stream) (void)hipStreamDestroy(s->stream);
    s->stream = nullptr;
}

// Fused spectral stage when it applies: tiled path, FFT radices 2-5, a y-z
// plane (two buffers) within the workgroup LDS limit (at most 144 KB); with
// the magnetic solve only for a symmetric stress (6 stress and 3
// polarisation moments) and without pairing.  The pencil group is the
// largest power of two up to 4 whose buffers fit in 64 KB.
template <typename T>
static int choose_spectral(ss_sim* s) {
    s->fused = false;
    if (!s->sorted || (s->p.flags & SS_FLAG_AXIS_PASSES) || s->fft.variant != SS_FFT_RADIX ||
        !s->fft.small_radices())
        return SS_OK;
    if (s->mag && (s->nsig != 6 || (s->p.flags & SS_FLAG_PAIRED_SPECTRAL))) return SS_OK;
    int lim = 0;
    SS_HIP_CHECK(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, s->p.device));
    const size_t cap = std::min<size_t>((size_t)lim, 144 * 1024);
    const SpecAxis a0 = spec_axis(s->fft.ax[0]), a1 = spec_axis(s->fft.ax[1]), a2 = spec_axis(s->fft.ax[2]);
    s->spec_p = 0;
    for (int P : {16, 32, 48, 64})
        if (ct_plan(a0, P) && ct_plan(a1, P) && ct_plan(a2, P)) s->spec_p = P;
    const size_t plane = spec_plane_lds(s->ny, s->nz, sizeof(T), s->spec_p > 0);
    const int nmax = std::max(s->nsig + (s->mag ? 3 : 0), s->nout);
    // four pencils (pencil pairs) per workgroup: at 64^3 more workgroups of
    // fewer lines finish sooner than fewer larger ones, down to this size
    int pb = 4;
    while (pb > 1 && spec_pencil_lds(s->nx, nmax * pb, sizeof(T)) > 64 * 1024) pb /= 2;
    const size_t pencil = spec_pencil_lds(s->nx, nmax * pb, sizeof(T));
    if (plane > cap || pencil > cap) {
        s->spec_p = 0;
        return SS_OK;
    }
    const bool pair = (s->p.flags & SS_FLAG_PAIRED_SPECTRAL) != 0;
    hipError_t ae = with_spec_p(s->spec_p, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        // the cap, not this simulation's size: the attribute belongs to the
        // kernel, which simulations of other sizes launch as well
        const int pl = (int)cap, pe = (int)cap;
        const hipFuncAttribute at = hipFuncAttributeMaxDynamicSharedMemorySize;
        hipError_t e = hipSuccess;
        auto set = [&](const void* f, int v) {
            if (e == hipSuccess) e = hipFuncSetAttribute(f, at, v);
        };
        if (pair) {
            set((const void*)k_spec_planes_fwd2<T, 6, P>, pl);
            set((const void*)k_spec_planes_fwd2<T, 9, P>, pl);
            set((const void*)k_spec_planes_inv2<T, P>, pl);
            set((const void*)k_spec_pencils2<T, P>, pe);
        } else {
            set((const void*)k_spec_planes_fwd<T, 6, P>, pl);
            set((const void*)k_spec_planes_fwd<T, 9, P>, pl);
            set((const void*)k_spec_planes_inv<T, P>, pl);
            set((const void*)k_spec_pencils<T, P>, pe);
        }
        return e;
    });
    SS_HIP_CHECK(ae);
    s->pencils = pb;
    s->fused = true;
    if (pair) {
        // pencil pairs (p, mirror p); a pair holds two pencils of half the
        // components, so the same pencil-group size fits the same LDS
        const int ny = s->ny, nz = s->nz, np = ny * nz;
        std::vector<uint2> pr;
        for (int p = 0; p < np; ++p) {
            const int y = p / nz, z = p % nz;
            const int m = ((ny - y) % ny) * nz + (nz - z) % nz;
            if (p <= m) pr.push_back(make_uint2((unsigned)p, (unsigned)m));
        }
        int rc = alloc(s, (void**)&s->pairs, pr.size() * sizeof(uint2));
        if (rc != SS_OK) return rc;
        SS_HIP_CHECK(hipMemcpy(s->pairs, pr.data(), pr.size() * sizeof(uint2), hipMemcpyHostToDevice));
        s->npairs = (int)pr.size();
        s->paired = true;
    }
    return SS_OK;
}

template <typename T>
static int create_impl(ss_sim* s) {
    const ss_params& p = s->p;
    int rc;
#define A(ptr, bytes)                                               \
    if ((rc = alloc(s, (void**)&(ptr), (bytes))) != SS_OK) return rc;
    const size_t n = s->n;
    A(s->x, n * sizeof(T));
    A(s->y, n * sizeof(T));
    A(s->z, n * sizeof(T));
    A(s->phi, n * sizeof(T));
    A(s->theta, n * sizeof(T));
    if (s->prec == SS_FP32) {
        A(s->lx, n * sizeof(lo_t));
        A(s->ly, n * sizeof(lo_t));
        A(s->lz, n * sizeof(lo_t));
    }
    if (s->sorted || s->psort) {
        A(s->x2, n * sizeof(T));
        A(s->y2, n * sizeof(T));
        A(s->z2, n * sizeof(T));
        A(s->phi2, n * sizeof(T));
        A(s->theta2, n * sizeof(T));
        if (s->prec == SS_FP32) {
            A(s->lx2, n * sizeof(lo_t));
            A(s->ly2, n * sizeof(lo_t));
            A(s->lz2, n * sizeof(lo_t));
        }
        A(s->id, n * 4);
        A(s->id2, n * 4);
        const int nt = s->tg.ntiles;
        A(s->tcnt, (size_t)nt * 4);
        A(s->toff, ((size_t)nt + 1) * 4);
        A(s->toff2, ((size_t)nt + 1) * 4);
        A(s->cursor, (size_t)nt * 4);
        if (s->sorted) {
            A(s->dcnt, (size_t)nt * NDC * 4);
            A(s->boff, (size_t)nt * NDC * 4);
            A(s->farcnt, (size_t)nt * 4);
            A(s->farbase, (size_t)nt * 4);
        }
        s->nparts = (nt + SCAN_CHUNK - 1) / SCAN_CHUNK;
        A(s->part, (size_t)s->nparts * 4);
    }
    if (s->psort) {
        A(s->strays, n * 4);
        A(s->scnt, (size_t)s->tg.ntiles * 4);
    }
    if (s->sorted) {
        A(s->mom_main, (size_t)s->ns * (s->nm + 1) * 8);
        A(s->mom_inflow, (size_t)s->ns * (s->nm + 1) * 8);
        SS_HIP_CHECK(hipMemset(s->mom_inflow, 0, (size_t)s->ns * (s->nm + 1) * 8));
        if (p.flags & SS_FLAG_KEEP_DISTRIBUTION) A(s->hist_next, (size_t)s->ns * s->na * 4);
    }
    if (!s->sorted || (p.flags & SS_FLAG_KEEP_DISTRIBUTION)) A(s->hist, (size_t)s->ns * s->na * 4);
    const size_t cs = 2 * sizeof(T) * (size_t)s->ns;
    A(s->spec_in, cs * s->nm);
    // (zeros in the components a step does not write: sigma_zz from the trace)
    SS_HIP_CHECK(hipMemset(s->spec_in, 0, cs * s->nm));
    A(s->spec_out, cs * s->nout);
    A(s->spec_full, cs * (12 + (s->mag ? 12 : 0)));
    if (s->use_rho) A(s->rho, (size_t)s->ns * sizeof(T));
    s->x_buf0 = s->x;
    s->toff_buf0 = s->toff;
    A(s->d_step, sizeof(uint64_t));
    SS_HIP_CHECK(hipMemset(s->d_step, 0, sizeof(uint64_t)));
    A(s->rec, (size_t)s->ns * s->rs * sizeof(T));
    SS_HIP_CHECK(hipMemset(s->rec, 0, (size_t)s->ns * s->rs * sizeof(T)));
    SS_HIP_CHECK(hipMemset(s->spec_in, 0, cs * s->nm));
    if (p.noise_mode == SS_NOISE_EXTERNAL) A(s->noise, 5 * n * sizeof(T));
    if (p.noise_mode == SS_NOISE_REFERENCE) A(s->noise, NOISE_REC * n * sizeof(T));
    if (p.noise_mode == SS_NOISE_REFERENCE && (s->sorted || s->psort)) A(s->inv, n * sizeof(uint32_t));
    const size_t stage_doubles = std::max<size_t>(5 * n, (size_t)24 * s->ns);
    A(s->staging, stage_doubles * sizeof(double));
    A(s->staging2, 3 * n * sizeof(double));
#undef A
    // tables
    std::vector<double> kxs, kys, kzs;
    k_sampling(s->nx, p.box_size.x, kxs);
    k_sampling(s->ny, p.box_size.y, kys);
    k_sampling(s->nz, p.box_size.z, kzs);
    std::vector<double> ik2(s->ns), kinv(s->ns);
    for (int i = 0; i < s->nx; ++i)
        for (int j = 0; j < s->ny; ++j)
            for (int k = 0; k < s->nz; ++k) {
                // get_inverse_norm_squared / get_inverse_norm (fft_helper.rs:79-100)
                double c = (kxs[i] * kxs[i] + kys[j] * kys[j]) + kzs[k] * kzs[k];
                size_t idx = ((size_t)i * s->ny + j) * s->nz + k;
                ik2[idx] = c / (c * c);
                kinv[idx] = 1.0 / sqrt(c);
            }
    ik2[0] = 0.0;
    kinv[0] = 0.0;
    if ((rc = upload<T>(&s->kx, kxs)) != SS_OK) return rc;
    s->owned.push_back(s->kx);
    if ((rc = upload<T>(&s->ky, kys)) != SS_OK) return rc;
    s->owned.push_back(s->ky);
    if ((rc = upload<T>(&s->kz, kzs)) != SS_OK) return rc;
    s->owned.push_back(s->kz);
    if ((rc = upload<T>(&s->ik2, ik2)) != SS_OK) return rc;
    s->owned.push_back(s->ik2);
    if ((rc = upload<T>(&s->kinv, kinv)) != SS_OK) return rc;
    s->owned.push_back(s->kinv);
    {
        std::vector<double> full;
        moment_kernels(p, 9, true, full);  // 9 stress components + orientation
        if ((rc = alloc(s, (void**)&s->diag_table, full.size() * 8)) != SS_OK) return rc;
        SS_HIP_CHECK(hipMemcpy(s->diag_table, full.data(), full.size() * 8, hipMemcpyHostToDevice));
        if ((rc = alloc(s, (void**)&s->diag_acc, DIAG_N * 8)) != SS_OK) return rc;
    }
    std::vector<double> kern;
    moment_kernels(p, s->nsig, s->mag, kern);
    if ((rc = upload<T>(&s->kern, kern)) != SS_OK) return rc;
    s->owned.push_back(s->kern);
    if (s->sorted) {
        // Fixed-point moment kernels: the largest possible cell sum,
        // N * max|K|, stays below 2^62.  The scale follows the largest kernel
        // value itself (small stress coefficients keep their relative
        // resolution); only all-zero kernels fall back to max|K| = 1.
        double kmax = 0.0;
        for (double v : kern) kmax = std::max(kmax, std::fabs(v));
        if (!(kmax > 0.0)) kmax = 1.0;
        const int shift = (int)std::floor(62.0 - std::log2((double)s->n * kmax));
        s->fixed_scale = std::ldexp(1.0, -shift);
        std::vector<long long> kf(kern.size());
        for (size_t i = 0; i < kern.size(); ++i) kf[i] = std::llround(std::ldexp(kern[i], shift));
        if ((rc = alloc(s, (void**)&s->kfix, kf.size() * 8)) != SS_OK) return rc;
        SS_HIP_CHECK(hipMemcpy(s->kfix, kf.data(), kf.size() * 8, hipMemcpyHostToDevice));
    }
    if ((rc = s->fft.init(s->nx, s->ny, s->nz, s->prec, p.fft_variant)) != SS_OK) return rc;
    s->dr = std::sqrt(2. * p.diffusion_rotational * p.timestep);
    if (p.noise_mode == SS_NOISE_REFERENCE &&
        (rc = refnoise_create(&s->refn, s->n, p.noise_streams, p.seed)) != SS_OK)
        return rc;
    rc = choose_spectral<T>(s);
    s->zz_trace = s->sorted && s->nm == 6 && !s->paired;
    return rc;
}

// Tile geometry of the tiled path.  Unless given, tiles start at 4^3 cells and
// shrink until there are at least 2048 of them (one workgroup each).  An axis
// whose length is at most tile + 1 is covered by one tile without halo.
static int choose_tiles(ss_sim* s) {
    const int n[3] = {s->nx, s->ny, s->nz};
    int t[3];
    const bool given = s->p.tile[0] > 0 && s->p.tile[1] > 0 && s->p.tile[2] > 0;
    for (int a = 0; a < 3; ++a) t[a] = given ? s->p.tile[a] : 4;
    auto count = [&]() {
        long long c = 1;
        for (int a = 0; a < 3; ++a) {
            int ta = std::min(t[a], n[a]);
            if (n[a] <= ta + 1) ta = n[a];
            c *= (n[a] + ta - 1) / ta;
        }
        return c;
    };
    while (!given && count() < 2048) {
        int best = -1;
        for (int a = 0; a < 3; ++a)
            if (t[a] > 1 && (best < 0 || t[a] > t[best])) best = a;
        if (best < 0) break;
        t[best] /= 2;
    }
    TileGeo& tg = s->tg;
    tg.ntiles = 1;
    tg.ws = 1;
    for (int a = 0; a < 3; ++a) {
        int ta = std::max(1, std::min(t[a], n[a]));
        if (n[a] <= ta + 1) ta = n[a];
        tg.t[a] = ta;
        tg.h[a] = ta == n[a] ? 0 : 1;
        tg.w[a] = ta + 2 * tg.h[a];
        tg.nt[a] = (n[a] + ta - 1) / ta;
        tg.ntiles *= tg.nt[a];
        tg.ws *= tg.w[a];
    }
    tg.nxcd = 8;
    // at least two steps: a re-sort takes its offsets from the direction
    // counts of a pass that did not re-sort
    s->resort = std::max(2, s->p.resort_interval > 0 ? s->p.resort_interval : 16);
    if (tile_lds_bytes(tg.ws, s->nm, s->na, s->nx + s->ny + s->nz) > (size_t)s->lds_max) {
        if (s->p.binning == SS_BIN_SORTED)
            return fail(SS_ERR_UNSUPPORTED, "tile window and angle bins exceed the LDS of the tiled binning");
        s->sorted = false;
    }
    return SS_OK;
}

// Tile geometry of the plain path with tile-ordered particles: tiles of
// 4 x 4 x 2 cells (clipped to the grid), halved along their longest axis until
// the tile's counts ([cells][angles] u32) fit 64 KiB of LDS (one cell may use
// up to the device's limit).  No halo.  False when even one cell does not
// fit, or when a count-grid index does not fit 32 bits (the drifted
// particles' list holds them).
static bool choose_plain_tiles(ss_sim* s) {
    if ((uint64_t)s->ns * (uint64_t)s->na >= (1ull << 32)) return false;
    const int n[3] = {s->nx, s->ny, s->nz};
    int t[3] = {std::min(4, n[0]), std::min(4, n[1]), std::min(2, n[2])};
    auto lds = [&]() { return (size_t)t[0] * t[1] * t[2] * s->na * 4; };
    while (lds() > 65536) {
        int best = -1;
        for (int a = 0; a < 3; ++a)
            if (t[a] > 1 && (best < 0 || t[a] > t[best])) best = a;
        if (best < 0) break;
        t[best] /= 2;
    }
    if (lds() > (size_t)s->lds_max) return false;
    TileGeo& tg = s->tg;
    tg.ntiles = 1;
    for (int a = 0; a < 3; ++a) {
        tg.t[a] = t[a];
        tg.h[a] = 0;
        tg.w[a] = t[a];
        tg.nt[a] = (n[a] + t[a] - 1) / t[a];
        tg.ntiles *= tg.nt[a];
    }
    tg.ws = t[0] * t[1] * t[2];
    tg.nxcd = 8;
    s->resort = s->p.resort_interval > 0 ? s->p.resort_interval : 16;
    s->psort_lds = (int)lds();
    return true;
}

static int validate(const ss_params* p) {
    if (!p) return fail(SS_ERR_INVALID, "params is NULL");
    const ss_grid& g = p->grid_size;
    if (g.x < 1 || g.y < 1 || g.z < 1 || g.phi < 1 || g.theta < 1)
        return fail(SS_ERR_INVALID, "grid sizes must be >= 1");
    if (!(p->box_size.x > 0) || !(p->box_size.y > 0) || !(p->box_size.z > 0))
        return fail(SS_ERR_INVALID, "box sizes must be > 0");
    if (p->number_of_particles < 1) return fail(SS_ERR_INVALID, "number_of_particles must be >= 1");
    if (p->number_of_particles >= (1ull << 31)) return fail(SS_ERR_UNSUPPORTED, "at most 2^31-1 particles");
    if (p->precision != SS_FP64 && p->precision != SS_FP32) return fail(SS_ERR_INVALID, "bad precision");
    if (p->noise_mode != SS_NOISE_PHILOX && p->noise_mode != SS_NOISE_EXTERNAL &&
        p->noise_mode != SS_NOISE_REFERENCE)
        return fail(SS_ERR_INVALID, "bad noise_mode");
    if (p->noise_mode == SS_NOISE_REFERENCE && p->noise_streams < 1)  // mod.rs:113-116
        return fail(SS_ERR_INVALID, "No environment variable 'RAYON_NUM_THREADS' set.");
    if (p->binning < SS_BIN_AUTO || p->binning > SS_BIN_PLAIN_TILED) return fail(SS_ERR_INVALID, "bad binning");
    if (p->fft_variant != SS_FFT_AUTO && p->fft_variant != SS_FFT_RADIX && p->fft_variant != SS_FFT_DFT)
        return fail(SS_ERR_INVALID, "bad fft_variant");
    if (p->fft_variant == SS_FFT_DFT && p->precision != SS_FP32)
        return fail(SS_ERR_UNSUPPORTED, "the dense-DFT FFT variant is FP32 only");
    const double ns = (double)g.x * g.y * g.z, na = (double)g.phi * g.theta;
    if (ns * na >= 4294967295.0) return fail(SS_ERR_UNSUPPORTED, "grid has 2^32 or more bins");
    if (na > 65535) return fail(SS_ERR_UNSUPPORTED, "at most 65535 angle bins");
    if ((p->flags & SS_FLAG_QUASI2D) && g.z != 1)
        return fail(SS_ERR_INVALID, "quasi2d needs grid_size.z == 1");
    return SS_OK;
}

}  // namespace ss

// ===========================================================================
// C ABI (every function below is declared extern "C" in include/ss_api.h)
// ===========================================================================

void ss_params_default(ss_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->box_size = {1.0, 1.0, 1.0};
    p->grid_size = {1, 1, 1, 1, 1};
    p->number_of_particles = 1;
    p->precision = SS_FP64;
    p->noise_mode = SS_NOISE_PHILOX;
    p->binning = SS_BIN_AUTO;
    p->fft_variant = SS_FFT_AUTO;
}

const char* ss_last_error(void) { return ss::g_err.c_str(); }

int ss_version(void) { return SS_API_VERSION; }

int ss_device_count(int* count) {
    if (!count) return fail(SS_ERR_INVALID, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return SS_OK;
}

int ss_create(const ss_params* params, ss_sim** out) {
    if (!out) return fail(SS_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int rc = validate(params);
    if (rc != SS_OK) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SS_ERR_DEVICE, "no GPU visible");
    if (params->device < 0 || params->device >= ndev) return fail(SS_ERR_INVALID, "bad device ordinal");
    SS_HIP_CHECK(hipSetDevice(params->device));
    ss_sim* s = new ss_sim();
    s->p = *params;
    const ss_params& p = s->p;
    {
        int lim = 0;
        if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, p.device) == hipSuccess && lim > 0)
            s->lds_max = lim;
    }
    s->prec = p.precision;
    s->esz = s->prec == SS_FP64 ? 8 : 4;
    s->sorted = p.binning == SS_BIN_SORTED || (p.binning == SS_BIN_AUTO && p.precision == SS_FP32);
    s->n = p.number_of_particles;
    // the tiled update addresses particle arrays through 32-bit byte offsets
    if (s->sorted && s->n > ((uint64_t)1 << 29)) {
        if (p.binning == SS_BIN_SORTED) {
            delete s;
            return fail(SS_ERR_UNSUPPORTED, "the tiled path holds at most 2^29 particles");
        }
        s->sorted = false;
    }
    s->nx = p.grid_size.x;
    s->ny = p.grid_size.y;
    s->nz = p.grid_size.z;
    s->np = p.grid_size.phi;
    s->nt = p.grid_size.theta;
    s->ns = s->nx * s->ny * s->nz;
    s->na = s->np * s->nt;
    s->use_e = p.shape != 0.0;
    s->use_rho = p.volume_exclusion != 0.0;
    s->use_graph = !(p.flags & (SS_FLAG_KEEP_DISTRIBUTION | SS_FLAG_NO_GRAPH));
    s->use_b = p.magnetic_dipole_dipole != 0.0;
    s->use_gb = p.magnetic_drag != 0.0;
    s->mag = s->use_b || s->use_gb || (p.flags & SS_FLAG_MAGNETIC_ALWAYS);
    s->nsig = p.stress_magnetic != 0.0 ? 9 : 6;
    s->nm = s->nsig + (s->mag ? 3 : 0);
    s->nout = 6 + (s->use_e ? 6 : 0) + (s->use_b ? 3 : 0) + (s->use_gb ? 9 : 0);
    s->rs = (s->use_e || s->use_rho || s->use_b || s->use_gb) ? REC_FULL : REC_SMALL;
    if (s->sorted) {
        int rc2 = choose_tiles(s);
        if (rc2 != SS_OK) {
            delete s;
            return rc2;
        }
    }
    if ((size_t)s->nm * s->na * s->esz > (size_t)s->lds_max) {
        delete s;
        return fail(SS_ERR_UNSUPPORTED, "too many angle bins for the LDS-resident moment kernels");
    }
    s->psort = !s->sorted && (p.binning == SS_BIN_PLAIN_TILED || (p.binning == SS_BIN_AUTO && p.precision == SS_FP64));
    if (s->psort && !choose_plain_tiles(s)) {
        if (p.binning == SS_BIN_PLAIN_TILED) {
            delete s;
            return fail(SS_ERR_UNSUPPORTED, "one cell's angle counts exceed the LDS of a workgroup, or the count "
                                            "grid has 2^32 bins or more");
        }
        s->psort = false;
    }
    // the end-of-period sort of this path is not captured
    if (s->psort) s->use_graph = false;
    if (hipStreamCreate(&s->stream) != hipSuccess) {
        delete s;
        return fail(SS_ERR_DEVICE, "hipStreamCreate failed");
    }
    rc = s->prec == SS_FP64 ? create_impl<double>(s) : create_impl<float>(s);
    if (rc != SS_OK) {
        std::string msg = ss_last_error();
        free_all(s);
        delete s;
        set_error(msg);
        return rc;
    }
    *out = s;
    return SS_OK;
}

int ss_destroy(ss_sim* s) {
    if (!s) return SS_OK;
    (void)hipSetDevice(s->p.device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_all(s);
    delete s;
    return SS_OK;
}

int ss_init(ss_sim* s, const double* particles, uint64_t n) {
    if (!s || !particles) return fail(SS_ERR_INVALID, "NULL argument");
    if (n != s->n)
        return fail(SS_ERR_INVALID, "Given initial condition has not the same number of particles (" +
                                        std::to_string(n) + ") as given in the parameter file (" +
                                        std::to_string(s->n) + ").");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    int rc = s->prec == SS_FP64 ? import_particles<double>(s, particles, nullptr, true)
                                : import_particles<float>(s, particles, nullptr, true);
    if (rc != SS_OK) return rc;
    SS_HIP_CHECK(hipStreamSynchronize(s->stream));
    s->timestep = 0;
    {
        const uint64_t z = 0;
        SS_HIP_CHECK(hipMemcpy(s->d_step, &z, sizeof(z), hipMemcpyHostToDevice));
    }
    s->initialised = true;
    s->have_fields = false;
    return SS_OK;
}

int ss_set_noise(ss_sim* s, const double* rv, uint64_t n) {
    if (!s || !rv) return fail(SS_ERR_INVALID, "NULL argument");
    if (s->p.noise_mode != SS_NOISE_EXTERNAL) return fail(SS_ERR_STATE, "noise_mode is not SS_NOISE_EXTERNAL");
    if (n != s->n) return fail(SS_ERR_INVALID, "noise length differs from number_of_particles");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    SS_HIP_CHECK(hipMemcpyAsync(s->staging, rv, sizeof(double) * 5 * n, hipMemcpyHostToDevice, s->stream));
    if (s->prec == SS_FP64)
        hipLaunchKernelGGL(k_import_noise<double>, dim3(blocks_for(n)), dim3(256), 0, s->stream, s->staging, n,
                           (double*)s->noise);
    else
        hipLaunchKernelGGL(k_import_noise<float>, dim3(blocks_for(n)), dim3(256), 0, s->stream, s->staging, n,
                           (float*)s->noise);
    SS_HIP_CHECK(hipGetLastError());
    SS_HIP_CHECK(hipStreamSynchronize(s->stream));
    s->have_noise = true;
    return SS_OK;
}

int ss_step(ss_sim* s, uint64_t nsteps, uint64_t* timestep_out) {
    if (!s) return fail(SS_ERR_INVALID, "NULL handle");
    if (!s->initialised) return fail(SS_ERR_STATE, "ss_step before ss_init");
    if (s->in_ensemble) return fail(SS_ERR_STATE, "an ensemble member steps with ss_ensemble_step");
    if (s->p.noise_mode == SS_NOISE_EXTERNAL && !s->have_noise)
        return fail(SS_ERR_STATE, "SS_NOISE_EXTERNAL needs ss_set_noise before ss_step");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    int rc = s->prec == SS_FP64 ? run_steps<double>(s, nsteps) : run_steps<float>(s, nsteps);
    if (rc != SS_OK) return rc;
    SS_HIP_CHECK(hipStreamSynchronize(s->stream));
    if (timestep_out) *timestep_out = s->timestep;
    return SS_OK;
}

int ss_prepare_steps(ss_sim* s, uint64_t nsteps) {
    if (!s) return fail(SS_ERR_INVALID, "NULL handle");
    if (!s->initialised) return fail(SS_ERR_STATE, "ss_prepare_steps before ss_init");
    if (s->in_ensemble) return fail(SS_ERR_STATE, "an ensemble member steps with ss_ensemble_step");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    return s->prec == SS_FP64 ? prepare_steps<double>(s, nsteps) : prepare_steps<float>(s, nsteps);
}

int ss_get_timestep(const ss_sim* s, uint64_t* t) {
    if (!s || !t) return fail(SS_ERR_INVALID, "NULL argument");
    *t = s->timestep;
    return SS_OK;
}

int ss_get_particles(ss_sim* s, double* out, uint64_t n) {
    if (!s || !out) return fail(SS_ERR_INVALID, "NULL argument");
    if (n > s->n) return fail(SS_ERR_INVALID, "n exceeds number_of_particles");
    if (!s->initialised) return fail(SS_ERR_STATE, "not initialised");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    return s->prec == SS_FP64 ? export_particles<double>(s, out, nullptr, n)
                              : export_particles<float>(s, out, nullptr, n);
}

int ss_get_distribution(ss_sim* s, uint32_t* counts, double* dist) {
    if (!s) return fail(SS_ERR_INVALID, "NULL handle");
    if (!s->initialised) return fail(SS_ERR_STATE, "not initialised");
    if (!s->have_hist || !s->hist)
        return fail(SS_ERR_STATE,
                    "no count grid kept: the sorted path needs SS_FLAG_KEEP_DISTRIBUTION or ss_keep_distribution");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    const size_t nb = (size_t)s->ns * s->na;
    std::vector<uint32_t> tmp;
    uint32_t* c = counts;
    if (!c) {
        tmp.resize(nb);
        c = tmp.data();
    }
    SS_HIP_CHECK(hipMemcpyAsync(c, s->hist, nb * sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream));
    SS_HIP_CHECK(hipStreamSynchronize(s->stream));
    if (dist) {
        // Distribution::sample_from then the volume rescale, in FP64 like the
        // reference (distribution/mod.rs:183, src/bin/simulation/mod.rs:255-257).
        GW gw = grid_width(s->p);
        double den = gw.x * gw.y * gw.z * gw.phi * gw.theta * (double)s->n;
        double vol = s->p.box_size.x * s->p.box_size.y * s->p.box_size.z;
        for (size_t i = 0; i < nb; ++i) dist[i] = ((double)c[i] / den) * vol;
    }
    return SS_OK;
}

int ss_get_flow_field(ss_sim* s, double* u, double* grad_u) {
    if (!s || !u) return fail(SS_ERR_INVALID, "NULL argument");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    const size_t ns = s->ns;
    if (!s->have_fields) {
        std::memset(u, 0, sizeof(double) * 3 * ns);
        if (grad_u) std::memset(grad_u, 0, sizeof(double) * 9 * ns);
        return SS_OK;
    }
    int rc = s->prec == SS_FP64 ? full_fields<double>(s) : full_fields<float>(s);
    if (rc != SS_OK) return rc;
    rc = s->prec == SS_FP64 ? copy_real<double>(s, s->spec_full, 0, 3, u) : copy_real<float>(s, s->spec_full, 0, 3, u);
    if (rc != SS_OK || !grad_u) return rc;
    return s->prec == SS_FP64 ? copy_real<double>(s, s->spec_full, 3, 9, grad_u)
                              : copy_real<float>(s, s->spec_full, 3, 9, grad_u);
}

int ss_get_magnetic_field(ss_sim* s, double* b, double* grad_b) {
    if (!s || !b) return fail(SS_ERR_INVALID, "NULL argument");
    if (!s->mag)
        return fail(SS_ERR_STATE,
                    "magnetic solve disabled: set SS_FLAG_MAGNETIC_ALWAYS or a nonzero magnetic coefficient");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    const size_t ns = s->ns;
    if (!s->have_fields) {
        std::memset(b, 0, sizeof(double) * 3 * ns);
        if (grad_b) std::memset(grad_b, 0, sizeof(double) * 9 * ns);
        return SS_OK;
    }
    int rc = s->prec == SS_FP64 ? full_fields<double>(s) : full_fields<float>(s);
    if (rc != SS_OK) return rc;
    rc = s->prec == SS_FP64 ? copy_real<double>(s, s->spec_full, 12, 3, b)
                            : copy_real<float>(s, s->spec_full, 12, 3, b);
    if (rc != SS_OK || !grad_b) return rc;
    return s->prec == SS_FP64 ? copy_real<double>(s, s->spec_full, 15, 9, grad_b)
                              : copy_real<float>(s, s->spec_full, 15, 9, grad_b);
}

int ss_get_snapshot(ss_sim* s, double* particles, double* compensation, uint64_t* timestep) {
    if (!s || !particles) return fail(SS_ERR_INVALID, "NULL argument");
    if (!s->initialised) return fail(SS_ERR_STATE, "not initialised");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    int rc = SS_OK;
    if (compensation)
        rc = s->prec == SS_FP64 ? export_particles<double>(s, particles, compensation, s->n)
                                : export_particles<float>(s, particles, compensation, s->n);
    else
        rc = s->prec == SS_FP64 ? export_particles<double>(s, particles, nullptr, s->n)
                                : export_particles<float>(s, particles, nullptr, s->n);
    if (timestep) *timestep = s->timestep;
    return rc;
}

int ss_resume(ss_sim* s, const double* particles, const double* compensation, uint64_t n, uint64_t timestep) {
    if (!s || !particles) return fail(SS_ERR_INVALID, "NULL argument");
    if (n != s->n) return fail(SS_ERR_INVALID, "snapshot size differs from number_of_particles");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    // Simulation::resume re-runs init (src/bin/simulation/mod.rs:199-207); the
    // snapshot's values are already in range, so they are taken as stored.
    int rc = s->prec == SS_FP64 ? import_particles<double>(s, particles, compensation, compensation == nullptr)
                                : import_particles<float>(s, particles, compensation, compensation == nullptr);
    if (rc != SS_OK) return rc;
    SS_HIP_CHECK(hipStreamSynchronize(s->stream));
    s->timestep = timestep;
    SS_HIP_CHECK(hipMemcpy(s->d_step, &timestep, sizeof(timestep), hipMemcpyHostToDevice));
    s->initialised = true;
    s->have_fields = false;
    return SS_OK;
}

int ss_get_rng_state(ss_sim* s, uint64_t* st, int32_t count) {
    if (!s) return fail(SS_ERR_INVALID, "NULL handle");
    if (!s->refn) return 0;
    if (!st) return fail(SS_ERR_INVALID, "NULL argument");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    return refnoise_get(s->refn, st, count, s->stream);
}

int ss_set_rng_state(ss_sim* s, const uint64_t* st, int32_t count) {
    if (!s) return fail(SS_ERR_INVALID, "NULL handle");
    if (!s->refn) return count == 0 ? SS_OK : fail(SS_ERR_STATE, "noise_mode has no generator state");
    if (count < 0 || (count > 0 && !st)) return fail(SS_ERR_INVALID, "bad rng state buffer");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    return refnoise_set(s->refn, st, count, s->stream);
}

int ss_get_noise(ss_sim* s, double* rv, uint64_t n) {
    if (!s || !rv) return fail(SS_ERR_INVALID, "NULL argument");
    if (s->p.noise_mode == SS_NOISE_PHILOX) return fail(SS_ERR_STATE, "Philox noise is drawn inside the update");
    if (n != s->n) return fail(SS_ERR_INVALID, "noise length differs from number_of_particles");
    if (!s->have_noise) return fail(SS_ERR_STATE, "no noise drawn yet");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    const bool recs = s->p.noise_mode == SS_NOISE_REFERENCE;
    const uint64_t stride = recs ? 1 : n;
    const int rec = recs ? NOISE_REC : 1;
    // the records sit at the slots of the order the last step's update read
    uint32_t* slots = nullptr;
    if (s->inv) {
        slots = reinterpret_cast<uint32_t*>(s->staging2);
        hipLaunchKernelGGL(k_inverse_perm, dim3(blocks_for(n)), dim3(256), 0, s->stream,
                           (const uint32_t*)(s->last_resort ? s->id2 : s->id), n, slots);
    }
    if (s->prec == SS_FP64)
        hipLaunchKernelGGL(k_export_noise<double>, dim3(blocks_for(n)), dim3(256), 0, s->stream,
                           (const double*)s->noise, n, stride, rec, (const uint32_t*)slots, s->staging);
    else
        hipLaunchKernelGGL(k_export_noise<float>, dim3(blocks_for(n)), dim3(256), 0, s->stream,
                           (const float*)s->noise, n, stride, rec, (const uint32_t*)slots, s->staging);
    SS_HIP_CHECK(hipGetLastError());
    SS_HIP_CHECK(hipMemcpyAsync(rv, s->staging, sizeof(double) * 5 * n, hipMemcpyDeviceToHost, s->stream));
    SS_HIP_CHECK(hipStreamSynchronize(s->stream));
    return SS_OK;
}

template <typename T>
static int diagnostics(ss_sim* s, double* out) {
    // |value| <= max(1, max|K|) per particle; N of them stay below 2^62
    double vmax = 1.0;
    {
        std::vector<double> full;
        moment_kernels(s->p, 9, true, full);
        for (double v : full) vmax = std::max(vmax, std::fabs(v));
    }
    const int shift = (int)std::floor(62.0 - std::log2((double)s->n * vmax));
    const double scale = std::ldexp(1.0, shift);
    SS_HIP_CHECK(hipMemsetAsync(s->diag_acc, 0, DIAG_N * 8, s->stream));
    // a grid-stride launch of eight workgroups per CU at most
    hipLaunchKernelGGL(k_diag<T>, dim3((unsigned)std::min<uint64_t>(blocks_for(s->n), 2048)), dim3(256), 0, s->stream,
                       state_a<T>(s), s->n, geo<T>(s),
                       (const double*)s->diag_table, scale, s->diag_acc);
    SS_HIP_CHECK(hipGetLastError());
    long long acc[DIAG_N];
    SS_HIP_CHECK(hipMemcpyAsync(acc, s->diag_acc, DIAG_N * 8, hipMemcpyDeviceToHost, s->stream));
    SS_HIP_CHECK(hipStreamSynchronize(s->stream));
    for (int q = 0; q < DIAG_N; ++q) out[q] = std::ldexp((double)acc[q], -shift) / (double)s->n;
    return SS_OK;
}

int ss_get_info(const ss_sim* s, ss_info* out) {
    if (!s || !out) return fail(SS_ERR_INVALID, "NULL argument");
    std::memset(out, 0, sizeof(*out));
    out->sorted = s->sorted ? 1 : 0;
    out->fused_spectral = s->fused ? (s->paired ? 2 : 1) : 0;
    out->plain_tiled = s->psort ? 1 : 0;
    if (s->sorted || s->psort) {
        for (int k = 0; k < 3; ++k) out->tile[k] = s->tg.t[k];
        out->ntiles = s->tg.ntiles;
        out->resort_interval = s->resort;
    }
    out->pencils = s->fused ? s->pencils : 0;
    out->graphs = s->use_graph ? 1 : 0;
    if (s->sorted) out->update_groups = (s->tg.ntiles + s->tg.nxcd - 1) / s->tg.nxcd * s->tg.nxcd;
    return SS_OK;
}

int ss_get_diagnostics(ss_sim* s, double* out) {
    if (!s || !out) return fail(SS_ERR_INVALID, "NULL argument");
    if (!s->initialised) return fail(SS_ERR_STATE, "not initialised");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    return s->prec == SS_FP64 ? diagnostics<double>(s, out) : diagnostics<float>(s, out);
}

int ss_synchronize(ss_sim* s) {
    if (!s) return fail(SS_ERR_INVALID, "NULL handle");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    SS_HIP_CHECK(hipStreamSynchronize(s->stream));
    return SS_OK;
}

int ss_keep_distribution(ss_sim* s, int on) {
    if (!s) return fail(SS_ERR_INVALID, "NULL handle");
    if (s->in_ensemble) return fail(SS_ERR_UNSUPPORTED, "ensemble members do not keep the distribution");
    if (!s->sorted) return SS_OK;
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    if (!on) {
        s->p.flags &= ~SS_FLAG_KEEP_DISTRIBUTION;
        s->use_graph = !(s->p.flags & SS_FLAG_NO_GRAPH);
        s->have_hist = false;
        s->swap_hist = false;
        return SS_OK;
    }
    const size_t nb = (size_t)s->ns * s->na;
    int rc;
    if (!s->hist && (rc = alloc(s, (void**)&s->hist, nb * 4)) != SS_OK) return rc;
    if (!s->hist_next && (rc = alloc(s, (void**)&s->hist_next, nb * 4)) != SS_OK) return rc;
    s->p.flags |= SS_FLAG_KEEP_DISTRIBUTION;
    drop_graphs(s);
    s->use_graph = false;
    if (s->initialised) {
        // the counts of the current particles: the distribution the next step samples
        fill_u32(s, s->hist, nb, 0u);
        if (s->prec == SS_FP64)
            hipLaunchKernelGGL(k_bin_hist<double>, dim3(blocks_for(s->n)), dim3(256), 0, s->stream, (const double*)s->x,
                               (const double*)s->y, (const double*)s->z, (const double*)s->phi,
                               (const double*)s->theta, s->n, geo<double>(s), s->hist);
        else
            hipLaunchKernelGGL(k_bin_hist<float>, dim3(blocks_for(s->n)), dim3(256), 0, s->stream, (const float*)s->x,
                               (const float*)s->y, (const float*)s->z, (const float*)s->phi, (const float*)s->theta,
                               s->n, geo<float>(s), s->hist);
        SS_HIP_CHECK(hipGetLastError());
        s->have_hist = true;
    }
    s->swap_hist = false;
    return SS_OK;
}

int ss_resort(ss_sim* s) {
    if (!s) return fail(SS_ERR_INVALID, "NULL handle");
    if (!s->initialised) return fail(SS_ERR_STATE, "ss_resort before ss_init");
    if (s->in_ensemble) return fail(SS_ERR_STATE, "ensemble members re-sort together with the ensemble");
    if (s->psort) {
        // the next step's binning re-sorts (already in tile order when
        // nothing moved since the last sort)
        if (s->steps_since_sort > 0) s->steps_since_sort = std::max<uint64_t>(s->steps_since_sort, s->resort);
        return SS_OK;
    }
    if (!s->sorted || s->pending) return SS_OK;
    if (s->last_resort) {
        // the last update re-sorted, so its direction counts do not describe
        // the new order: re-sort after the next step instead
        s->steps_since_sort = std::max<uint64_t>(s->steps_since_sort, (uint64_t)s->resort - 1);
        return SS_OK;
    }
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    // offsets from the counts binned by the last update (or by ss_init); the
    // next step's update writes the particles in the new tile order
    hipError_t e = launch_count_scan(s);
    if (e != hipSuccess) return fail(SS_ERR_DEVICE, std::string("resort: ") + hipGetErrorString(e));
    return SS_OK;
}

int ss_profile_steps(ss_sim* s, uint64_t nsteps, double* ms_per_step, char* names, int name_stride,
                     int max_stages) {
    if (!s || !ms_per_step) return fail(SS_ERR_INVALID, "NULL argument");
    if (!s->initialised) return fail(SS_ERR_STATE, "ss_profile_steps before ss_init");
    if (s->in_ensemble) return fail(SS_ERR_STATE, "an ensemble member steps with the ensemble");
    if (s->p.noise_mode == SS_NOISE_EXTERNAL && !s->have_noise)
        return fail(SS_ERR_STATE, "SS_NOISE_EXTERNAL needs ss_set_noise");
    if (nsteps == 0) return fail(SS_ERR_INVALID, "nsteps must be >= 1");
    SS_HIP_CHECK(hipSetDevice(s->p.device));
    hipEvent_t ev[ST_COUNT + 1];
    for (int i = 0; i <= ST_COUNT; ++i) SS_HIP_CHECK(hipEventCreate(&ev[i]));
    double acc[ST_COUNT] = {0};
    int rc = SS_OK;
    for (uint64_t k = 0; k < nsteps && rc == SS_OK; ++k) {
        rc = s->prec == SS_FP64 ? step_once<double>(s, ev) : step_once<float>(s, ev);
        if (rc != SS_OK) break;
        if (hipEventSynchronize(ev[ST_COUNT]) != hipSuccess) {
            rc = fail(SS_ERR_DEVICE, "event sync failed");
            break;
        }
        for (int st = 0; st < ST_COUNT; ++st) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[st], ev[st + 1]);
            acc[st] += ms;
        }
    }
    for (int i = 0; i <= ST_COUNT; ++i) (void)hipEventDestroy(ev[i]);
    if (rc != SS_OK) return rc;
    const int nst = std::min(max_stages, (int)ST_COUNT);
    for (int st = 0; st < nst; ++st) {
        ms_per_step[st] = acc[st] / (double)nsteps;
        if (names && name_stride > 0) {
            std::strncpy(names + (size_t)st * name_stride, STAGE_NAMES[st], name_stride - 1);
            names[(size_t)st * name_stride + name_stride - 1] = 0;
        }
    }
    return nst;
}

// ---------------------------------------------------------------------------
// Solver entry points (unit-level parity).
// ---------------------------------------------------------------------------

// A planned solve on given distributions: a one-particle system on the plain
// path owns the transforms, tables and spectra, plus a device copy of the
// distribution.  Created once, reused for every solve.
struct ss_solver {
    ss_sim* s = nullptr;
    double* d = nullptr;  // [Ns][Na] distribution
    int kind = SS_SOLVER_FLOW;
};


template <typename T>
static int solver_run(ss_solver* v, const double* dist, double* o1, double* o2) {
    ss_sim* s = v->s;
    const size_t nb = (size_t)s->ns * s->na;
    SS_HIP_CHECK(hipMemcpyAsync(v->d, dist, nb * sizeof(double), hipMemcpyHostToDevice, s->stream));
    Geo<T> g = geo<T>(s);
    GW gw = grid_width(s->p);
    const T measure = (T)(gw.phi * gw.theta);
    const unsigned nbk = (unsigned)((s->ns + 3) / 4);
    if (s->nm == 6)
        hipLaunchKernelGGL((k_contract_dense<T, 6>), dim3(nbk), dim3(256), 0, s->stream, v->d, g, (const T*)s->kern,
                           measure, (cplx<T>*)s->spec_in);
    else if (s->nm == 9)
        hipLaunchKernelGGL((k_contract_dense<T, 9>), dim3(nbk), dim3(256), 0, s->stream, v->d, g, (const T*)s->kern,
                           measure, (cplx<T>*)s->spec_in);
    else
        hipLaunchKernelGGL((k_contract_dense<T, 12>), dim3(nbk), dim3(256), 0, s->stream, v->d, g,
                           (const T*)s->kern, measure, (cplx<T>*)s->spec_in);
    SS_HIP_CHECK(hipGetLastError());
    int rc = s->fft.run(s->spec_in, s->nm, -1, s->stream);
    if (rc == SS_OK) rc = full_fields<T>(s);
    if (rc == SS_OK) {
        const int c0 = v->kind == SS_SOLVER_MAGNETIC ? 12 : 0;
        rc = copy_real<T>(s, s->spec_full, c0, 3, o1);
        if (rc == SS_OK && o2) rc = copy_real<T>(s, s->spec_full, c0 + 3, 9, o2);
    }
    return rc;
}


int ss_solver_create(const ss_params* params, int kind, ss_solver** out) {
    if (!params || !out) return fail(SS_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (kind != SS_SOLVER_FLOW && kind != SS_SOLVER_MAGNETIC) return fail(SS_ERR_INVALID, "bad solver kind");
    ss_params p = *params;
    p.number_of_particles = 1;
    p.noise_mode = SS_NOISE_PHILOX;
    p.binning = SS_BIN_HISTOGRAM;
    if (kind == SS_SOLVER_MAGNETIC) p.flags |= SS_FLAG_MAGNETIC_ALWAYS;
    p.flags &= ~SS_FLAG_QUASI2D;
    ss_solver* v = new ss_solver();
    v->kind = kind;
    int rc = ss_create(&p, &v->s);
    if (rc != SS_OK) {
        delete v;
        return rc;
    }
    const size_t nb = (size_t)v->s->ns * v->s->na;
    if (hipMalloc(&v->d, nb * sizeof(double)) != hipSuccess) {
        ss_destroy(v->s);
        delete v;
        return fail(SS_ERR_ALLOC, "hipMalloc");
    }
    *out = v;
    return SS_OK;
}

int ss_solver_destroy(ss_solver* v) {
    if (!v) return SS_OK;
    if (v->s) (void)hipSetDevice(v->s->p.device);
    if (v->d) (void)hipFree(v->d);
    ss_destroy(v->s);
    delete v;
    return SS_OK;
}

int ss_solver_mean_flow_field(ss_solver* v, double hydro_screening, const double* dist, double* u, double* grad_u) {
    if (!v || !dist || !u) return fail(SS_ERR_INVALID, "NULL argument");
    if (v->kind != SS_SOLVER_FLOW) return fail(SS_ERR_INVALID, "not a flow solver");
    SS_HIP_CHECK(hipSetDevice(v->s->p.device));
    v->s->p.hydro_screening = hydro_screening;  // read by the solve at launch
    return v->s->prec == SS_FP64 ? solver_run<double>(v, dist, u, grad_u) : solver_run<float>(v, dist, u, grad_u);
}

int ss_solver_mean_magnetic_field(ss_solver* v, const double* dist, double* b, double* grad_b) {
    if (!v || !dist || !b) return fail(SS_ERR_INVALID, "NULL argument");
    if (v->kind != SS_SOLVER_MAGNETIC) return fail(SS_ERR_INVALID, "not a magnetic solver");
    SS_HIP_CHECK(hipSetDevice(v->s->p.device));
    return v->s->prec == SS_FP64 ? solver_run<double>(v, dist, b, grad_b) : solver_run<float>(v, dist, b, grad_b);
}

static int solve_once(const ss_params* params, int kind, const double* dist, double* o1, double* o2) {
    ss_solver* v = nullptr;
    int rc = ss_solver_create(params, kind, &v);
    if (rc != SS_OK) return rc;
    rc = kind == SS_SOLVER_FLOW ? ss_solver_mean_flow_field(v, params->hydro_screening, dist, o1, o2)
                                : ss_solver_mean_magnetic_field(v, dist, o1, o2);
    std::string msg = ss_last_error();
    ss_solver_destroy(v);
    if (rc != SS_OK) set_error(msg);
    return rc;
}

int ss_flow_solve(const ss_params* params, const double* dist, double* u, double* grad_u) {
    if (!params || !dist || !u) return fail(SS_ERR_INVALID, "NULL argument");
    return solve_once(params, SS_SOLVER_FLOW, dist, u, grad_u);
}

int ss_magnetic_solve(const ss_params* params, const double* dist, double* b, double* grad_b) {
    if (!params || !dist || !b) return fail(SS_ERR_INVALID, "NULL argument");
    return solve_once(params, SS_SOLVER_MAGNETIC, dist, b, grad_b);
}


// DensityGradient::get_real_gradient (distribution/density_gradient.rs).
template <typename T>
static int density_gradient(const ss_params* params, const double* dist, double* grad) {
    ss_params p = *params;
    p.number_of_particles = 1;
    p.noise_mode = SS_NOISE_PHILOX;
    p.binning = SS_BIN_HISTOGRAM;
    p.flags &= ~SS_FLAG_QUASI2D;
    ss_sim* s = nullptr;
    int rc = ss_create(&p, &s);
    if (rc != SS_OK) return rc;
    const size_t nb = (size_t)s->ns * s->na;
    double *d = nullptr, *rho = nullptr;
    if (hipMalloc(&d, nb * sizeof(double)) != hipSuccess || hipMalloc(&rho, (size_t)s->ns * sizeof(double)) != hipSuccess)
        rc = fail(SS_ERR_ALLOC, "hipMalloc");
    if (rc == SS_OK && hipMemcpy(d, dist, nb * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(SS_ERR_DEVICE, "copy in");
    if (rc == SS_OK) {
        GW gw = grid_width(s->p);
        const double dthph = gw.theta * gw.phi;
        const dim3 grid(blocks_for((uint64_t)s->ns));
        hipLaunchKernelGGL(k_density_rows, grid, dim3(256), 0, s->stream, d, (uint64_t)s->ns, s->na, dthph, rho);
        hipLaunchKernelGGL(k_density_spec<T>, grid, dim3(256), 0, s->stream, rho, (uint64_t)s->ns, (T)dthph,
                           (cplx<T>*)s->spec_in);
        if (hipGetLastError() != hipSuccess) rc = fail(SS_ERR_DEVICE, "density launch");
    }
    if (rc == SS_OK) rc = s->fft.run(s->spec_in, 1, -1, s->stream);
    if (rc == SS_OK) {
        const T nre = T(0) / (T)s->ns, nim = T(1) / (T)s->ns;  // Complex(0, 1) / n
        hipLaunchKernelGGL(k_density_grad<T>, dim3(blocks_for((uint64_t)s->ns)), dim3(256), 0, s->stream,
                           (const cplx<T>*)s->spec_in, geo<T>(s), (const T*)s->kx, (const T*)s->ky,
                           (const T*)s->kz, nre, nim, (cplx<T>*)s->spec_out);
        if (hipGetLastError() != hipSuccess) rc = fail(SS_ERR_DEVICE, "gradient launch");
    }
    if (rc == SS_OK) rc = s->fft.run(s->spec_out, 3, +1, s->stream);
    if (rc == SS_OK) rc = copy_real<T>(s, s->spec_out, 0, 3, grad);
    (void)hipFree(d);
    (void)hipFree(rho);
    std::string msg = ss_last_error();
    ss_destroy(s);
    if (rc != SS_OK) set_error(msg);
    return rc;
}

int ss_density_gradient(const ss_params* params, const double* dist, double* grad) {
    if (!params || !dist || !grad) return fail(SS_ERR_INVALID, "NULL argument");
    return params->precision == SS_FP64 ? density_gradient<double>(params, dist, grad)
                                        : density_gradient<float>(params, dist, grad);
}

int ss_interpolate_vector_field(const ss_params* params, const double* field, const double* positions, uint64_t n,
                                double* out) {
    if (!params || !field || (n && (!positions || !out))) return fail(SS_ERR_INVALID, "NULL argument");
    const ss_grid& g = params->grid_size;
    if (g.x < 1 || g.y < 1 || g.z < 1) return fail(SS_ERR_INVALID, "grid sizes must be >= 1");
    if (n == 0) return SS_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SS_ERR_DEVICE, "no GPU visible");
    if (params->device < 0 || params->device >= ndev) return fail(SS_ERR_INVALID, "bad device ordinal");
    SS_HIP_CHECK(hipSetDevice(params->device));
    const size_t nf = 3 * (size_t)g.x * g.y * g.z;
    double *df = nullptr, *dp = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc(&df, nf * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dp, 3 * n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dout, 3 * n * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(df, field, nf * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dp, positions, 3 * n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        GW gw = grid_width(*params);
        hipLaunchKernelGGL(k_interpolate, dim3(blocks_for(n)), dim3(256), 0, 0, df, g.x, g.y, g.z, gw.x, gw.y, gw.z,
                           dp, n, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, 3 * n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(df);
    (void)hipFree(dp);
    (void)hipFree(dout);
    if (e != hipSuccess) return fail(SS_ERR_DEVICE, std::string("interpolation: ") + hipGetErrorString(e));
    return SS_OK;
}

template <typename T>
static int debug_update(const ss_params* p, uint64_t n, const double* parts, const double* fields,
                        const double* rv, double* out) {
    ss_sim tmp;
    tmp.p = *p;
    tmp.use_e = true;
    tmp.use_rho = true;
    tmp.use_b = true;
    tmp.use_gb = true;
    UpdateParams<T> P = update_params<T>(&tmp);
    double *dp = nullptr, *df = nullptr, *dr = nullptr, *dout = nullptr;
    int rc = SS_OK;
    if (hipMalloc(&dp, 5 * n * 8) != hipSuccess || hipMalloc(&df, 34 * n * 8) != hipSuccess ||
        hipMalloc(&dr, 5 * n * 8) != hipSuccess || hipMalloc(&dout, 5 * n * 8) != hipSuccess)
        rc = fail(SS_ERR_ALLOC, "hipMalloc");
    if (rc == SS_OK && (hipMemcpy(dp, parts, 5 * n * 8, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(df, fields, 34 * n * 8, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(dr, rv, 5 * n * 8, hipMemcpyHostToDevice) != hipSuccess))
        rc = fail(SS_ERR_DEVICE, "copy in");
    if (rc == SS_OK) {
        hipLaunchKernelGGL(k_debug_update<T>, dim3(blocks_for(n)), dim3(256), 0, 0, dp, df, dr, n, P,
                           (T)p->magnetic_dipole_dipole, dout);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
            rc = fail(SS_ERR_DEVICE, "debug update failed");
    }
    if (rc == SS_OK && hipMemcpy(out, dout, 5 * n * 8, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(SS_ERR_DEVICE, "copy out");
    (void)hipFree(dp);
    (void)hipFree(df);
    (void)hipFree(dr);
    (void)hipFree(dout);
    return rc;
}

int ss_debug_update(const ss_params* params, uint64_t n, const double* particles, const double* fields,
                    const double* rv, double* out) {
    if (!params || !particles || !fields || !rv || !out) return fail(SS_ERR_INVALID, "NULL argument");
    if (n == 0) return SS_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SS_ERR_DEVICE, "no GPU visible");
    SS_HIP_CHECK(hipSetDevice(params->device));
    return params->precision == SS_FP64 ? debug_update<double>(params, n, particles, fields, rv, out)
                                        : debug_update<float>(params, n, particles, fields, rv, out);
}

template <typename T>
static int debug_c2g(const ss_params* p, uint64_t n, const double* parts, int32_t* out) {
    ss_sim tmp;
    tmp.p = *p;
    tmp.nx = p->grid_size.x;
    tmp.ny = p->grid_size.y;
    tmp.nz = p->grid_size.z;
    tmp.np = p->grid_size.phi;
    tmp.nt = p->grid_size.theta;
    tmp.ns = tmp.nx * tmp.ny * tmp.nz;
    tmp.na = tmp.np * tmp.nt;
    Geo<T> g = geo<T>(&tmp);
    double* dp = nullptr;
    int32_t* dout = nullptr;
    int rc = SS_OK;
    if (hipMalloc(&dp, 5 * n * 8) != hipSuccess || hipMalloc(&dout, 5 * n * 4) != hipSuccess)
        rc = fail(SS_ERR_ALLOC, "hipMalloc");
    if (rc == SS_OK && hipMemcpy(dp, parts, 5 * n * 8, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(SS_ERR_DEVICE, "copy in");
    if (rc == SS_OK) {
        hipLaunchKernelGGL(k_coord_to_grid<T>, dim3(blocks_for(n)), dim3(256), 0, 0, dp, n, g, dout);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
            rc = fail(SS_ERR_DEVICE, "coord_to_grid failed");
    }
    if (rc == SS_OK && hipMemcpy(out, dout, 5 * n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(SS_ERR_DEVICE, "copy out");
    (void)hipFree(dp);
    (void)hipFree(dout);
    return rc;
}

int ss_debug_coord_to_grid(const ss_params* params, uint64_t n, const double* particles, int32_t* out) {
    if (!params || !particles || !out) return fail(SS_ERR_INVALID, "NULL argument");
    if (n == 0) return SS_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(SS_ERR_DEVICE, "no GPU visible");
    SS_HIP_CHECK(hipSetDevice(params->device));
    return params->precision == SS_FP64 ? debug_c2g<double>(params, n, particles, out)
                                        : debug_c2g<float>(params, n, particles, out);
}

// ---------------------------------------------------------------------------
// Initial condition (host): uniform positions; cos(theta) by inverse CDF of
// 1 + a sin(2 pi x / Lx) cos(theta); phi uniform.
// ---------------------------------------------------------------------------
int ss_initial_condition(const ss_params* params, double amplitude, uint64_t seed, double* particles,
                         uint64_t n) {
    if (!params || !particles) return fail(SS_ERR_INVALID, "NULL argument");
    const double L[3] = {params->box_size.x, params->box_size.y, params->box_size.z};
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t w[8];
        philox4x32_10((uint32_t)i, 0xC0FFEEu, (uint32_t)(i >> 32), 0u, k0, k1, w);
        philox4x32_10((uint32_t)i, 0xC0FFEEu, (uint32_t)(i >> 32), 1u, k0, k1, w + 4);
        double u[5];
        for (int k = 0; k < 5; ++k) u[k] = ((double)w[k] + 0.5) * (1.0 / 4294967296.0);
        double* p = particles + 5 * i;
        p[0] = L[0] * u[0];
        p[1] = L[1] * u[1];
        p[2] = L[2] * u[2];
        p[3] = TWOPI_D * u[3];
        const double eps = amplitude * sin(TWOPI_D * p[0] / L[0]);
        // density (1 + eps c)/2 on c in [-1,1]; CDF F(c) = ((c+1) + eps (c^2-1)/2)/2
        const double q = 1.0 - eps / 2.0 - 2.0 * u[4];  // eps c^2/2 + c + q = 0
        double c = -2.0 * q / (1.0 + sqrt(1.0 - 2.0 * eps * q));
        c = std::min(1.0, std::max(-1.0, c));
        p[4] = acos(c);
    }
    return SS_OK;
}

// ===========================================================================
// Batched ensemble: independent systems stepped together.  The per-system
// bin conversion, FFTs, k-space solve and packing run on NSIDE side streams
// (parallel branches of the captured graph); the particle update of all
// systems is one persistent launch over the concatenated tile ranges.
// ===========================================================================
static const int NSIDE = 4;
// pencils per workgroup of a batched spectral stage (when the LDS allows):
// 8 against the single system's 4 took the 48-system benchmark ensemble from
// 2.51 to 2.47 ms per step; 16 made it 2.59 ms
static const int ENS_PENCILS = 8;

struct ss_ensemble {
    std::vector<ss_sim*> m;
    hipStream_t stream = nullptr;
    hipStream_t side[NSIDE] = {nullptr};
    hipEvent_t fork = nullptr, join[NSIDE] = {nullptr};
    uint64_t* d_step = nullptr;
    int prec = SS_FP32;
    int nm = 6;
    int resort = 16;
    bool ext = false;
    bool full = true;  // some member needs the terms the default-physics update compiles out
    // persistent-update job tables: [slot 0/1][in place / re-sort] -> nsys TileArgs
    void* d_jobs = nullptr;
    int* d_start = nullptr;
    size_t job_bytes = 0;
    std::vector<int> built_for;  // member parities slot 0 was built for
    int total_blocks = 0;
    size_t lds = 0;
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};  // [pending][slot]
    uint64_t timestep = 0;
    uint64_t steps_since_sort = 0;
    // batched spectral stage: members on the fused per-component path, grouped
    // by compile-time axis length; each group is one launch of each kernel
    struct SpecGroup {
        int p = 0;                 // spec_p of the members
        std::vector<int> members;  // indices into m
        int blocks[3] = {0, 0, 0}; // planes forward, pencils, planes inverse
        size_t lds_plane = 0, lds_pencil = 0;
        size_t job_off = 0;        // byte offset of its SpecJob array in d_specjobs
        size_t first_off = 0;      // int offset of its three first[] tables in d_specfirst
        int side = 0;              // side stream of its launches
    };
    std::vector<SpecGroup> sgroups;
    std::vector<char> batched;  // per member: its spectral stage runs in a group
    void* d_specjobs = nullptr;
    int* d_specfirst = nullptr;
};

namespace ss {

static int member_parity(const ss_sim* s) { return s->x == s->x_buf0 ? 0 : 1; }

// Job tables for the members' current buffers (slot 0) and the swapped ones
// (slot 1), each for an update in place and for an update that carries out
// a pending re-sort.
// The in-place persistent update that ens_launch_update runs for an
// ensemble of this noise mode and physics.
template <typename T, int NM>
static const void* ens_update_kernel(bool ext, bool full) {
    if (ext) return (const void*)k_tile_update_batched<T, NM, true, true, false>;
    if (NM == 6 && !full) return (const void*)k_tile_update_batched<T, NM, true, NZ_PHILOX, false, false>;
    return (const void*)k_tile_update_batched<T, NM, true, false, false>;
}

// Resident 256-thread workgroups per CU of a kernel with `lds` bytes of LDS
// each: the lower of the LDS bound and the register bound (a workgroup puts
// one wave on each of the CU's four SIMDs; a SIMD holds 512 VGPRs per lane
// and at most eight waves).
static int workgroups_per_cu(const void* kern, size_t lds, int device) {
    int by_regs = 8;
    hipFuncAttributes fa{};
    if (hipFuncGetAttributes(&fa, kern) == hipSuccess && fa.numRegs > 0)
        by_regs = std::min(8, 512 / ((fa.numRegs + 7) / 8 * 8));
    int lds_cu = 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess ||
        lds_cu <= 0)
        lds_cu = 65536;
    return std::min(by_regs, lds > 0 ? (int)((size_t)lds_cu / lds) : by_regs);
}

template <typename T, int NM>
static int ens_build_jobs(ss_ensemble* e) {
    const int n = (int)e->m.size();
    std::vector<TileArgs<T, NM>> h((size_t)4 * n);
    for (int slot = 0; slot < 2; ++slot) {
        for (ss_sim* s : e->m)
            if (slot == 1) {  // the swapped configuration, restored below
                swap_ab(s);
                std::swap(s->toff, s->toff2);
            }
        for (int sc = 0; sc < 2; ++sc)
            for (int j = 0; j < n; ++j) {
                ss_sim* s = e->m[j];
                const bool pend = s->pending;
                s->pending = sc == 1;
                h[((size_t)slot * 2 + sc) * n + j] = tile_args<T, NM>(s);
                s->pending = pend;
            }
        for (ss_sim* s : e->m)
            if (slot == 1) {
                swap_ab(s);
                std::swap(s->toff, s->toff2);
            }
    }
    SS_HIP_CHECK(hipMemcpy(e->d_jobs, h.data(), h.size() * sizeof(TileArgs<T, NM>), hipMemcpyHostToDevice));
    e->full = false;
    for (int j = 0; j < n; ++j) e->full = e->full || full_physics(h[j].P);
    e->built_for.clear();
    for (ss_sim* s : e->m) e->built_for.push_back(member_parity(s));
    return SS_OK;
}

template <typename T, int NM>
static void ens_launch_update(ss_ensemble* e, int slot, bool resort, int k) {
    const int n = (int)e->m.size();
    const TileArgs<T, NM>* jobs =
        reinterpret_cast<const TileArgs<T, NM>*>(e->d_jobs) + ((size_t)slot * 2 + (resort ? 1 : 0)) * n;
    const int grid = std::max(1, std::min(e->total_blocks, 256 * 8));
    const size_t lds = e->lds + sizeof(int) * (n + 1);
    const int nj = (int)e->lds;
    const uint32_t ko = (uint32_t)k;
    auto go = [&](auto kern) {
        allow_lds(e->m[0]->lds_max, (const void*)kern, lds);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, e->stream, jobs, e->d_start, n, nj, ko);
    };
    if (e->ext) {
        if (resort)
            go(k_tile_update_batched<T, NM, true, true, true>);
        else
            go(k_tile_update_batched<T, NM, true, true, false>);
    } else if (NM == 6 && !e->full) {  // every member on the default physics
        if (resort)
            go(k_tile_update_batched<T, NM, true, NZ_PHILOX, true, false>);
        else
            go(k_tile_update_batched<T, NM, true, NZ_PHILOX, false, false>);
    } else {
        if (resort)
            go(k_tile_update_batched<T, NM, true, false, true>);
        else
            go(k_tile_update_batched<T, NM, true, false, false>);
    }
}

// Groups of the batched spectral stage: every member on the fused
// per-component path joins the group of its compile-time axis length (moment
// layouts of 6 or 9 components); the others keep their own launches.
template <typename T>
static int ens_build_spec(ss_ensemble* e) {
    const int n = (int)e->m.size();
    e->batched.assign(n, 0);
    e->sgroups.clear();
    if (e->nm != 6 && e->nm != 9) return SS_OK;
    for (int j = 0; j < n; ++j) {
        const ss_sim* s = e->m[j];
        if (!s->fused || s->paired || s->mag) continue;  // the batched kernels have no magnetic solve
        ss_ensemble::SpecGroup* g = nullptr;
        for (auto& q : e->sgroups)
            if (q.p == s->spec_p) g = &q;
        if (!g) {
            e->sgroups.emplace_back();
            g = &e->sgroups.back();
            g->p = s->spec_p;
        }
        g->members.push_back(j);
        e->batched[j] = 1;
    }
    // Balance the side streams: a group with more cells than a stream's even
    // share is split into chunks of consecutive members of about equal cells,
    // and the groups go to the streams largest first, each to the stream with
    // the fewest cells so far.
    {
        auto cells = [&](int j) { return (double)e->m[j]->nx * e->m[j]->ny * e->m[j]->nz; };
        auto work = [&](const ss_ensemble::SpecGroup& g) {
            double w = 0;
            for (int j : g.members) w += cells(j);
            return w;
        };
        double total = 0;
        for (const auto& g : e->sgroups) total += work(g);
        std::vector<ss_ensemble::SpecGroup> split;
        for (const auto& g : e->sgroups) {
            const double w = work(g);
            const int k = std::max(1, std::min((int)g.members.size(), (int)std::lround(w * NSIDE / total)));
            double acc = 0;
            int c = 0;
            split.emplace_back();
            split.back().p = g.p;
            for (int j : g.members) {
                if (c + 1 < k && acc >= (c + 1) * w / k) {
                    ++c;
                    split.emplace_back();
                    split.back().p = g.p;
                }
                split.back().members.push_back(j);
                acc += cells(j);
            }
        }
        std::vector<size_t> order(split.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](size_t a, size_t b) { return work(split[a]) > work(split[b]); });
        double load[NSIDE] = {0};
        for (size_t i : order) {
            int q = 0;
            for (int r = 1; r < NSIDE; ++r)
                if (load[r] < load[q]) q = r;
            split[i].side = q;
            load[q] += work(split[i]);
        }
        e->sgroups.swap(split);
    }
    if (e->sgroups.empty()) return SS_OK;
    int lim = 0;
    SS_HIP_CHECK(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, e->m[0]->p.device));
    const int cap = std::min(lim, 144 * 1024);
    std::vector<SpecJob<T>> jobs;
    std::vector<int> first;
    for (auto& g : e->sgroups) {
        g.job_off = jobs.size() * sizeof(SpecJob<T>);
        g.first_off = first.size();
        std::vector<int> f[3];
        for (int k = 0; k < 3; ++k) f[k].assign(1, 0);
        for (int j : g.members) {
            ss_sim* s = e->m[j];
            double ms, ds;
            scales(s, ms, ds);
            SpecJob<T> a;
            a.main = s->mom_main;
            a.inflow = s->mom_inflow;
            a.spec_in = (cplx<T>*)s->spec_in;
            a.spec_out = (cplx<T>*)s->spec_out;
            a.rho = s->use_rho ? (T*)s->rho : nullptr;
            a.rec = (T*)s->rec;
            a.twx = (const cplx<T>*)s->fft.ax[0].tw;
            a.twy = (const cplx<T>*)s->fft.ax[1].tw;
            a.twz = (const cplx<T>*)s->fft.ax[2].tw;
            a.kx = (const T*)s->kx;
            a.ky = (const T*)s->ky;
            a.kz = (const T*)s->kz;
            a.ik2 = (const T*)s->ik2;
            a.kinv = (const T*)s->kinv;
            a.scale = s->fixed_scale * ms;
            a.rho_scale = ds;
            a.screening = (T)s->p.hydro_screening;
            a.norm = (T)(double)s->ns;
            a.ax = spec_axis(s->fft.ax[0]);
            a.ay = spec_axis(s->fft.ax[1]);
            a.az = spec_axis(s->fft.ax[2]);
            a.nx = s->nx;
            a.np = s->ny * s->nz;
            a.nz = s->nz;
            a.rs = s->rs;
            a.nsig = s->nsig;
            a.nout = s->use_e ? 12 : 6;
            a.zz_trace = s->zz_trace ? 1 : 0;
            // a batched launch has workgroups to spare: more pencils per
            // workgroup make the x-line reads longer runs of each row
            a.pb = s->pencils;
            if (ENS_PENCILS > a.pb && a.np >= ENS_PENCILS &&
                spec_pencil_lds(s->nx, std::max(a.nsig, a.nout) * ENS_PENCILS, sizeof(T)) <= (size_t)cap)
                a.pb = ENS_PENCILS;
            jobs.push_back(a);
            // the block counts of launch_spec_planes_fwd / _pencils / _planes_inv
            f[0].push_back(f[0].back() + s->nx * (s->nm + 1));
            f[1].push_back(f[1].back() + (a.np + a.pb - 1) / a.pb);
            f[2].push_back(f[2].back() + s->nx * a.nout);
            g.lds_plane = std::max(g.lds_plane, spec_plane_lds(s->ny, s->nz, sizeof(T), s->spec_p > 0));
            g.lds_pencil = std::max(g.lds_pencil, spec_pencil_lds(s->nx, std::max(a.nsig, a.nout) * a.pb, sizeof(T)));
        }
        for (int k = 0; k < 3; ++k) {
            g.blocks[k] = f[k].back();
            first.insert(first.end(), f[k].begin(), f[k].end());
        }
    }
    for (auto& g : e->sgroups) {
        if (g.lds_plane > (size_t)cap || g.lds_pencil > (size_t)cap)
            return fail(SS_ERR_UNSUPPORTED, "batched spectral stage: LDS of a group exceeds the device limit");
        hipError_t err = with_spec_p(g.p, [&](auto pc) {
            constexpr int P = decltype(pc)::value;
            const hipFuncAttribute at = hipFuncAttributeMaxDynamicSharedMemorySize;
            hipError_t r = hipFuncSetAttribute((const void*)k_spec_planes_fwd_batched<T, 6, P>, at, cap);
            if (r == hipSuccess) r = hipFuncSetAttribute((const void*)k_spec_planes_fwd_batched<T, 9, P>, at, cap);
            if (r == hipSuccess) r = hipFuncSetAttribute((const void*)k_spec_pencils_batched<T, P>, at, cap);
            if (r == hipSuccess) r = hipFuncSetAttribute((const void*)k_spec_planes_inv_batched<T, P>, at, cap);
            return r;
        });
        SS_HIP_CHECK(err);
    }
    if (hipMalloc(&e->d_specjobs, jobs.size() * sizeof(SpecJob<T>)) != hipSuccess ||
        hipMalloc((void**)&e->d_specfirst, first.size() * sizeof(int)) != hipSuccess)
        return fail(SS_ERR_ALLOC, "hipMalloc");
    SS_HIP_CHECK(hipMemcpy(e->d_specjobs, jobs.data(), jobs.size() * sizeof(SpecJob<T>), hipMemcpyHostToDevice));
    SS_HIP_CHECK(hipMemcpy(e->d_specfirst, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice));
    return SS_OK;
}

// The spectral stage of group g on stream st: three launches.
template <typename T>
static hipError_t ens_launch_spec(ss_ensemble* e, const ss_ensemble::SpecGroup& g, hipStream_t st) {
    const int n = (int)g.members.size();
    const SpecJob<T>* jobs = reinterpret_cast<const SpecJob<T>*>((const char*)e->d_specjobs + g.job_off);
    const int* f0 = e->d_specfirst + g.first_off;
    const int* f1 = f0 + (n + 1);
    const int* f2 = f1 + (n + 1);
    return with_spec_p(g.p, [&](auto pc) {
        constexpr int P = decltype(pc)::value;
        if (e->nm == 6)
            hipLaunchKernelGGL((k_spec_planes_fwd_batched<T, 6, P>), dim3(g.blocks[0]), dim3(SPEC_PLANE_THREADS),
                               g.lds_plane, st, jobs, f0, n);
        else
            hipLaunchKernelGGL((k_spec_planes_fwd_batched<T, 9, P>), dim3(g.blocks[0]), dim3(SPEC_PLANE_THREADS),
                               g.lds_plane, st, jobs, f0, n);
        hipLaunchKernelGGL((k_spec_pencils_batched<T, P>), dim3(g.blocks[1]), dim3(SPEC_PENCIL_THREADS),
                           g.lds_pencil, st, jobs, f1, n);
        hipLaunchKernelGGL((k_spec_planes_inv_batched<T, P>), dim3(g.blocks[2]), dim3(SPEC_PLANE_THREADS),
                           g.lds_plane, st, jobs, f2, n);
        return hipGetLastError();
    });
}

// Job-table slot matching the members' current buffers.
static int ens_slot(const ss_ensemble* e) { return member_parity(e->m[0]) == e->built_for[0] ? 0 : 1; }

// One ensemble step; `k` is the step's index within the re-sort period (the
// step counter offset while capturing).
template <typename T>
static int ens_step_once(ss_ensemble* e, int k, bool capturing, hipEvent_t* ev = nullptr) {
    hipError_t err;
    if (ev) SS_HIP_CHECK(hipEventRecord(ev[0], e->stream));
    SS_HIP_CHECK(hipEventRecord(e->fork, e->stream));
    for (int q = 0; q < NSIDE; ++q) SS_HIP_CHECK(hipStreamWaitEvent(e->side[q], e->fork, 0));
    for (size_t g = 0; g < e->sgroups.size(); ++g) {
        if ((err = ens_launch_spec<T>(e, e->sgroups[g], e->side[e->sgroups[g].side])) != hipSuccess)
            return fail(SS_ERR_DEVICE, std::string("batched spectral stage: ") + hipGetErrorString(err));
    }
    for (size_t j = 0; j < e->m.size(); ++j) {
        ss_sim* s = e->m[j];
        if (e->batched[j]) {  // what spectral_stage records for the fused path
            s->spec_partial = true;
            s->spec_packed = false;
            continue;
        }
        int rc = spectral_stage<T>(s, nullptr);
        if (rc != SS_OK) return rc;
    }
    for (int q = 0; q < NSIDE; ++q) {
        SS_HIP_CHECK(hipEventRecord(e->join[q], e->side[q]));
        SS_HIP_CHECK(hipStreamWaitEvent(e->stream, e->join[q], 0));
    }
    if (ev) SS_HIP_CHECK(hipEventRecord(ev[1], e->stream));
    const int slot = ens_slot(e);
    const bool resort = e->m[0]->pending;
    if (e->nm == 6)
        ens_launch_update<T, 6>(e, slot, resort, k);
    else if (e->nm == 9)
        ens_launch_update<T, 9>(e, slot, resort, k);
    else
        ens_launch_update<T, 12>(e, slot, resort, k);
    SS_HIP_CHECK(hipGetLastError());
    if (ev) SS_HIP_CHECK(hipEventRecord(ev[2], e->stream));
    e->steps_since_sort += 1;
    for (ss_sim* s : e->m) {
        if (resort) swap_sorted(s);
        s->steps_since_sort += 1;
        s->timestep += 1;
        s->have_fields = true;
    }
    if (e->steps_since_sort >= (uint64_t)e->resort) {
        // the next re-sort's offsets, on the side streams, then join
        SS_HIP_CHECK(hipEventRecord(e->fork, e->stream));
        for (int q = 0; q < NSIDE; ++q) SS_HIP_CHECK(hipStreamWaitEvent(e->side[q], e->fork, 0));
        for (ss_sim* s : e->m)
            if ((err = launch_count_scan(s)) != hipSuccess) return fail(SS_ERR_DEVICE, hipGetErrorString(err));
        for (int q = 0; q < NSIDE; ++q) {
            SS_HIP_CHECK(hipEventRecord(e->join[q], e->side[q]));
            SS_HIP_CHECK(hipStreamWaitEvent(e->stream, e->join[q], 0));
        }
        e->steps_since_sort = 0;
    }
    if (ev) SS_HIP_CHECK(hipEventRecord(ev[3], e->stream));
    if (!capturing) hipLaunchKernelGGL(k_add_step, dim3(1), dim3(64), 0, e->stream, e->d_step, 1u);
    e->timestep += 1;
    return SS_OK;
}

template <typename T>
static int ens_jobs_for_current(ss_ensemble* e) {
    bool same = e->built_for.size() == e->m.size();
    for (size_t j = 0; same && j < e->m.size(); ++j) same = e->built_for[j] == member_parity(e->m[j]);
    if (same) return 0;
    bool flipped = e->built_for.size() == e->m.size();
    for (size_t j = 0; flipped && j < e->m.size(); ++j) flipped = e->built_for[j] != member_parity(e->m[j]);
    if (flipped) return 1;
    for (int k = 0; k < 4; ++k)
        if (e->graph[k]) {
            destroy_exec(e->stream, e->graph[k]);
            e->graph[k] = nullptr;
        }
    int rc = e->nm == 6 ? ens_build_jobs<T, 6>(e) : e->nm == 9 ? ens_build_jobs<T, 9>(e) : ens_build_jobs<T, 12>(e);
    return rc == SS_OK ? 0 : rc;
}

template <typename T>
static int ens_run(ss_ensemble* e, uint64_t nsteps) {
    const int K = e->resort;
    while (nsteps > 0) {
        const int slot = ens_jobs_for_current<T>(e);
        if (slot < 0) return slot;
        if (e->steps_since_sort == 0 && nsteps >= (uint64_t)K) {
            const bool pending = e->m[0]->pending;
            const int key = (pending ? 2 : 0) + slot;
            if (!e->graph[key]) {
                SS_HIP_CHECK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
                int rc = SS_OK;
                for (int k = 0; k < K && rc == SS_OK; ++k) rc = ens_step_once<T>(e, k, true);
                if (rc == SS_OK)
                    hipLaunchKernelGGL(k_add_step, dim3(1), dim3(64), 0, e->stream, e->d_step, (uint32_t)K);
                hipGraph_t g = nullptr;
                hipError_t err = hipStreamEndCapture(e->stream, &g);
                if (rc != SS_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (err != hipSuccess) return fail(SS_ERR_DEVICE, std::string("ensemble capture: ") + hipGetErrorString(err));
                err = hipGraphInstantiate(&e->graph[key], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (err != hipSuccess) {
                    e->graph[key] = nullptr;
                    return fail(SS_ERR_DEVICE, std::string("ensemble instantiate: ") + hipGetErrorString(err));
                }
            } else {
                // advance the bookkeeping the captured period performs
                for (ss_sim* s : e->m) {
                    s->timestep += K;
                    if (pending) swap_sorted(s);
                    s->pending = true;
                    s->steps_since_sort = 0;
                    s->have_fields = true;
                    s->spec_partial = s->fused;
                    s->spec_packed = s->fused && s->paired;
                }
                e->timestep += K;
            }
            SS_HIP_CHECK(hipGraphLaunch(e->graph[key], e->stream));
            nsteps -= (uint64_t)K;
        } else {
            // eager step: the job table's offsets are relative to the counter
            int rc = ens_step_once<T>(e, 0, false);
            if (rc != SS_OK) return rc;
            nsteps -= 1;
        }
    }
    return SS_OK;
}

}  // namespace ss

int ss_ensemble_create(const ss_params* params, int nsys, ss_ensemble** out) {
    if (!params || !out) return fail(SS_ERR_INVALID, "NULL argument");
    *out = nullptr;
    if (nsys < 1 || nsys > MAX_BATCH) return fail(SS_ERR_INVALID, "nsys must be in [1, 256]");
    ss_ensemble* e = new ss_ensemble();
    auto bail = [&](int rc) {
        std::string msg = ss_last_error();
        ss_ensemble_destroy(e);
        set_error(msg);
        return rc;
    };
    for (int j = 0; j < nsys; ++j) {
        ss_params p = params[j];
        p.binning = SS_BIN_SORTED;
        p.flags &= ~(SS_FLAG_KEEP_DISTRIBUTION);
        p.flags |= SS_FLAG_NO_GRAPH;
        if (p.precision != params[0].precision || p.noise_mode != params[0].noise_mode || p.device != params[0].device)
            return bail(fail(SS_ERR_INVALID, "ensemble members must share precision, noise mode and device"));
        if (p.noise_mode == SS_NOISE_REFERENCE)
            return bail(fail(SS_ERR_UNSUPPORTED, "ensembles draw Philox or external noise"));
        if (p.resort_interval != params[0].resort_interval)
            return bail(fail(SS_ERR_INVALID, "ensemble members must share the re-sort interval"));
        ss_sim* s = nullptr;
        int rc = ss_create(&p, &s);
        if (rc != SS_OK) return bail(rc);
        e->m.push_back(s);
        if (!s->sorted) return bail(fail(SS_ERR_UNSUPPORTED, "member " + std::to_string(j) + " cannot use the tiled path"));
        if (s->nm != e->m[0]->nm) return bail(fail(SS_ERR_INVALID, "ensemble members must share the moment layout"));
    }
    ss_sim* s0 = e->m[0];
    e->prec = s0->prec;
    e->nm = s0->nm;
    e->resort = s0->resort;
    e->ext = s0->p.noise_mode == SS_NOISE_EXTERNAL;
    if (hipSetDevice(s0->p.device) != hipSuccess) return bail(fail(SS_ERR_DEVICE, "hipSetDevice"));
    if (hipStreamCreate(&e->stream) != hipSuccess || hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess)
        return bail(fail(SS_ERR_DEVICE, "stream/event creation failed"));
    for (int q = 0; q < NSIDE; ++q)
        if (hipStreamCreate(&e->side[q]) != hipSuccess ||
            hipEventCreateWithFlags(&e->join[q], hipEventDisableTiming) != hipSuccess)
            return bail(fail(SS_ERR_DEVICE, "stream/event creation failed"));
    if (hipMalloc(&e->d_step, sizeof(uint64_t)) != hipSuccess || hipMemset(e->d_step, 0, sizeof(uint64_t)) != hipSuccess)
        return bail(fail(SS_ERR_ALLOC, "hipMalloc"));
    // members run on the side streams and count steps on the shared counter
    std::vector<int> start(1, 0);
    size_t lds = 0, lds_staged = 0;
    for (int j = 0; j < nsys; ++j) {
        ss_sim* s = e->m[j];
        (void)hipStreamDestroy(s->stream);
        s->stream = e->side[j % NSIDE];
        s->in_ensemble = true;
        s->d_step = e->d_step;
        const int per = (s->tg.ntiles + s->tg.nxcd - 1) / s->tg.nxcd;
        start.push_back(start.back() + per * s->tg.nxcd);
        lds = std::max(lds, tile_lds_bytes(s->tg.ws, s->nm, s->na, s->nx + s->ny + s->nz));
        lds_staged = std::max(lds_staged, tile_lds_bytes(s->tg.ws, s->nm, s->na, s->nx + s->ny + s->nz, tile_rec_bytes(s)));
    }
    // staged cell records only when they leave room for the member table and
    // keep as many workgroups per CU: the persistent update sizes its LDS for
    // the largest window of any member, and a workgroup fewer per CU costs
    // more than the staging saves (48 members of 32^3 to 64^3 cells: 2.34 ms
    // staged at two workgroups per CU against 1.88 ms unstaged at three)
    {
        bool full = false;
        for (ss_sim* s : e->m)
            full = full || (e->prec == SS_FP32 ? full_physics(update_params<float>(s)) : full_physics(update_params<double>(s)));
        const void* kern = e->prec == SS_FP32
                               ? (e->nm == 6 ? ens_update_kernel<float, 6>(e->ext, full)
                                             : e->nm == 9 ? ens_update_kernel<float, 9>(e->ext, full)
                                                          : ens_update_kernel<float, 12>(e->ext, full))
                               : (e->nm == 6 ? ens_update_kernel<double, 6>(e->ext, full)
                                             : e->nm == 9 ? ens_update_kernel<double, 9>(e->ext, full)
                                                          : ens_update_kernel<double, 12>(e->ext, full));
        const size_t table = sizeof(int) * (nsys + 1);
        const size_t with = (lds_staged + 15) / 16 * 16 + table, without = (lds + 15) / 16 * 16 + table;
        if (with <= (size_t)e->m[0]->lds_max &&
            workgroups_per_cu(kern, with, s0->p.device) >= workgroups_per_cu(kern, without, s0->p.device))
            lds = lds_staged;
        else
            for (ss_sim* s : e->m) s->no_stage = true;
    }
    e->total_blocks = start.back();
    e->lds = (lds + 15) / 16 * 16;
    // the persistent update keeps the member table beside the largest window
    if (e->lds + sizeof(int) * (nsys + 1) > (size_t)e->m[0]->lds_max)
        return bail(fail(SS_ERR_UNSUPPORTED,
                         "ensemble: the largest member tile window plus the member table exceed the LDS of a "
                         "workgroup; use smaller tiles"));
    e->job_bytes = e->prec == SS_FP32 ? (e->nm == 6 ? sizeof(TileArgs<float, 6>) : e->nm == 9 ? sizeof(TileArgs<float, 9>) : sizeof(TileArgs<float, 12>))
                                      : (e->nm == 6 ? sizeof(TileArgs<double, 6>) : e->nm == 9 ? sizeof(TileArgs<double, 9>) : sizeof(TileArgs<double, 12>));
    if (hipMalloc(&e->d_jobs, e->job_bytes * 4 * nsys) != hipSuccess ||
        hipMalloc(&e->d_start, sizeof(int) * (nsys + 1)) != hipSuccess)
        return bail(fail(SS_ERR_ALLOC, "hipMalloc"));
    if (hipMemcpy(e->d_start, start.data(), sizeof(int) * (nsys + 1), hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(SS_ERR_DEVICE, "copy"));
    {
        const int rc = e->prec == SS_FP64 ? ens_build_spec<double>(e) : ens_build_spec<float>(e);
        if (rc != SS_OK) return bail(rc);
    }
    *out = e;
    return SS_OK;
}

int ss_ensemble_destroy(ss_ensemble* e) {
    if (!e) return SS_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (int k = 0; k < 4; ++k)
        if (e->graph[k]) (void)hipGraphExecDestroy(e->graph[k]);
    for (ss_sim* s : e->m) {
        if (s->d_step == e->d_step) s->d_step = nullptr;
        bool shared = false;
        for (int q = 0; q < NSIDE; ++q) shared = shared || s->stream == e->side[q];
        if (shared) s->stream = nullptr;
        ss_destroy(s);
    }
    for (int q = 0; q < NSIDE; ++q) {
        if (e->side[q]) (void)hipStreamDestroy(e->side[q]);
        if (e->join[q]) (void)hipEventDestroy(e->join[q]);
    }
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    if (e->d_step) (void)hipFree(e->d_step);
    if (e->d_jobs) (void)hipFree(e->d_jobs);
    if (e->d_start) (void)hipFree(e->d_start);
    if (e->d_specjobs) (void)hipFree(e->d_specjobs);
    if (e->d_specfirst) (void)hipFree(e->d_specfirst);
    delete e;
    return SS_OK;
}

int ss_ensemble_profile_steps(ss_ensemble* e, uint64_t nsteps, double* ms_per_step, char* names, int name_stride,
                              int max_stages) {
    static const char* NAMES[3] = {"spectral", "update", "sort"};
    if (!e || !ms_per_step) return fail(SS_ERR_INVALID, "NULL argument");
    if (nsteps == 0) return fail(SS_ERR_INVALID, "nsteps must be >= 1");
    for (ss_sim* s : e->m) {
        if (!s->initialised) return fail(SS_ERR_STATE, "every member needs ss_init before profiling");
        if (s->p.noise_mode == SS_NOISE_EXTERNAL && !s->have_noise)
            return fail(SS_ERR_STATE, "SS_NOISE_EXTERNAL needs ss_set_noise for every member");
    }
    SS_HIP_CHECK(hipSetDevice(e->m[0]->p.device));
    e->timestep = e->m[0]->timestep;
    e->steps_since_sort = e->m[0]->steps_since_sort;
    for (int q = 0; q < NSIDE; ++q) SS_HIP_CHECK(hipStreamSynchronize(e->side[q]));
    SS_HIP_CHECK(hipMemcpy(e->d_step, &e->timestep, sizeof(uint64_t), hipMemcpyHostToDevice));
    hipEvent_t ev[4];
    for (int i = 0; i < 4; ++i) SS_HIP_CHECK(hipEventCreate(&ev[i]));
    double acc[3] = {0.0, 0.0, 0.0};
    int rc = SS_OK;
    for (uint64_t k = 0; k < nsteps && rc == SS_OK; ++k) {
        const int slot = e->prec == SS_FP64 ? ens_jobs_for_current<double>(e) : ens_jobs_for_current<float>(e);
        if (slot < 0) {
            rc = slot;
            break;
        }
        rc = e->prec == SS_FP64 ? ens_step_once<double>(e, 0, false, ev) : ens_step_once<float>(e, 0, false, ev);
        if (rc != SS_OK) break;
        if (hipEventSynchronize(ev[3]) != hipSuccess) {
            rc = fail(SS_ERR_DEVICE, "event sync failed");
            break;
        }
        for (int st = 0; st < 3; ++st) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[st], ev[st + 1]);
            acc[st] += ms;
        }
    }
    for (int i = 0; i < 4; ++i) (void)hipEventDestroy(ev[i]);
    if (rc != SS_OK) return rc;
    SS_HIP_CHECK(hipStreamSynchronize(e->stream));
    const int nst = std::min(max_stages, 3);
    for (int st = 0; st < nst; ++st) {
        ms_per_step[st] = acc[st] / (double)nsteps;
        if (names && name_stride > 0) {
            std::strncpy(names + (size_t)st * name_stride, NAMES[st], name_stride - 1);
            names[(size_t)st * name_stride + name_stride - 1] = 0;
        }
    }
    return nst;
}

int ss_ensemble_member(ss_ensemble* e, int sys, ss_sim** out) {
    if (!e || !out) return fail(SS_ERR_INVALID, "NULL argument");
    if (sys < 0 || sys >= (int)e->m.size()) return fail(SS_ERR_INVALID, "member index out of range");
    *out = e->m[sys];
    return SS_OK;
}

int ss_ensemble_step(ss_ensemble* e, uint64_t nsteps, uint64_t* timestep_out) {
    if (!e) return fail(SS_ERR_INVALID, "NULL handle");
    for (ss_sim* s : e->m) {
        if (!s->initialised) return fail(SS_ERR_STATE, "every member needs ss_init before ss_ensemble_step");
        if (s->p.noise_mode == SS_NOISE_EXTERNAL && !s->have_noise)
            return fail(SS_ERR_STATE, "SS_NOISE_EXTERNAL needs ss_set_noise for every member");
        if (s->timestep != e->m[0]->timestep || s->steps_since_sort != e->m[0]->steps_since_sort ||
            s->pending != e->m[0]->pending)
            return fail(SS_ERR_STATE, "ensemble members are out of step: initialise them together");
    }
    SS_HIP_CHECK(hipSetDevice(e->m[0]->p.device));
    e->timestep = e->m[0]->timestep;
    e->steps_since_sort = e->m[0]->steps_since_sort;
    // members' own eager work (init, getters) ran on the side streams
    for (int q = 0; q < NSIDE; ++q) SS_HIP_CHECK(hipStreamSynchronize(e->side[q]));
    SS_HIP_CHECK(hipMemcpy(e->d_step, &e->timestep, sizeof(uint64_t), hipMemcpyHostToDevice));
    int rc = e->prec == SS_FP64 ? ens_run<double>(e, nsteps) : ens_run<float>(e, nsteps);
    if (rc != SS_OK) return rc;
    SS_HIP_CHECK(hipStreamSynchronize(e->stream));
    if (timestep_out) *timestep_out = e->timestep;
    return SS_OK;
}
