// cns_api.hip -- host side of the C ABI (include/cns.h): configuration,
// grid and metric assembly, device buffers, the step loop and diagnostics.
//
// Reference counterparts:
//   cns_create        setGPUParameters + initSolver (src/cuda_utils.cu:49-179,415-519)
//   grid / metrics    initGrid, derivGrid (src/init.cpp:32-92,258-277)
//   cns_set_state     copyField(0) (src/cuda_utils.cu:299-331)
//   cns_get_state     copyField(1) (src/cuda_utils.cu:332-355)
//   cns_rhs           calcRHS (src/cuda_main.cu:15-42)
//   cns_step          runSimulation / runSimulationLowStorage (src/cuda_main.cu:44-186)
//   dt, forcing       calcTimeStepPressGrad (src/cuda_main.cu:249-265), kept on the device
//   cns_destroy       clearSolver (src/cuda_utils.cu:521)
//   diagnostics       calcAvgChan, printRes (src/init.cpp:150-256) on the device
//   field files       writeField / initField (src/init.cpp:13-30), saveFileMPI /
//                     readFileMPI (src/comm.cpp:205-279), grid files (src/init.cpp:78-88)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/cns.h"
#include "cns_aux.hpp"
#include "cns_batch.hpp"
#include "cns_diag.hpp"
#include "cns_launch.hpp"

using namespace cns;

namespace {

thread_local std::string g_err;  // errors without a handle

// exact rational central-difference weights (order 2s), reference layout:
// cF[it] multiplies f[i+it-s]-f[i+s-it]; cS[it] multiplies f[i+it-s]+f[i+s-it],
// cS[s] is the centre weight (src/globals.h:69-95)
int64_t igcd(int64_t a, int64_t b) {
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) {
        int64_t t = a % b;
        a = b;
        b = t;
    }
    return a ? a : 1;
}
int64_t ifact(int n) {
    int64_t f = 1;
    for (int i = 2; i <= n; i++) f *= i;
    return f;
}
void fd_weights(int s, double *cF, double *cS) {
    const int64_t sf2 = ifact(s) * ifact(s);
    int64_t cn = 0, cd = 1;
    for (int k = 1; k <= s; k++) {
        int64_t num = sf2, den = (int64_t)k * ifact(s - k) * ifact(s + k);
        int64_t g = igcd(num, den);
        num /= g;
        den /= g;
        const bool odd = (k % 2) == 1;
        cF[s - k] = (odd ? -(double)num : (double)num) / (double)den;
        int64_t n2 = 2 * num, d2 = den * k;
        g = igcd(n2, d2);
        n2 /= g;
        d2 /= g;
        cS[s - k] = (odd ? (double)n2 : -(double)n2) / (double)d2;
        const int64_t sn = odd ? n2 : -n2;
        int64_t nn = cn * d2 + sn * cd, dn = cd * d2;
        g = igcd(nn, dn);
        cn = nn / g;
        cd = dn / g;
    }
    int64_t nc = -2 * cn, dc = cd;
    const int64_t g = igcd(nc, dc);
    nc /= g;
    dc /= g;
    cS[s] = (double)nc / (double)dc;
}

// SplitMix64 and its uniform double in [0,1): (x >> 11) * 2^-53
struct SplitMix64 {
    uint64_t x;
    uint64_t next() {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Split [0, n) over the host's hardware threads (at most 16) for the
// per-point host loops of initial conditions and FP32 packing.
template <class F>
void host_parallel(long n, F f) {
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt < 1 ? 1 : (nt > 16 ? 16 : nt);
    if (n < 65536 || nt == 1) {
        f(0L, n);
        return;
    }
    std::vector<std::thread> th;
    const long chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        const long a = t * chunk, b = std::min(n, a + chunk);
        if (a < b) th.emplace_back([=]() { f(a, b); });
    }
    for (auto &x : th) x.join();
}

}  // namespace

struct cns_solver {
    cns_config cfg;
    std::string err;
    int S = 0, SV = 0;
    long n = 0;
    int dev = 0;
    hipStream_t stream = nullptr;
    // grid (host)
    std::vector<double> x, y, z, xp, xpp, dxv, vsx;
    double dx = 0, d_dx = 0, d_dy = 0, d_dz = 0, d_d2x = 0, d_d2y = 0, d_d2z = 0, Rgas = 0, Ec = 0;
    double cF[4] = {0}, cS[5] = {0}, cVF[4] = {0}, cVS[5] = {0};
    // device
    double *d_xp = nullptr, *d_vsx = nullptr, *d_dxv = nullptr, *d_spx = nullptr, *d_spz = nullptr;
    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr;  // coordinates (device initial conditions)
    double *d_ref = nullptr;  // 5 sponge reference fields, mx*mz each
    Scalars *d_sc = nullptr;
    StepCtl *d_ctl = nullptr;  // per-member step control when advanced in a batch
    bool in_batch = false;
    long bsteps = 0;  // steps taken in batches since the last set_state / cns_step (CFL phase)
    double *d_part = nullptr;
    // state storage: FP64, or FP32 offsets from (rho0, E0) with optional
    // compensation (low) parts
    bool fp32 = false, comp = false;
    size_t esz = 8;
    double rho0 = 0.0, E0 = 0.0;
    void *d_dil = nullptr;
    int nsets = 0, nrhs = 0;
    void *sets[4] = {nullptr};  // conserved-field sets, 5*n each
    void *lo[4] = {nullptr};    // FP32 compensation parts of the sets
    void *rbuf[2] = {nullptr};  // right-hand-side sets, 5*n each
    int cur = 0;
    double *d_par1 = nullptr, *d_par2 = nullptr, *d_time = nullptr;
    double *d_prof = nullptr;  // 10*mx wall-normal profiles (allocated on first use)
    double *d_stats = nullptr;  // flow statistics samples of cns_step_stats
    int stats_cap = 0;
    int par_cap = 0;
    bool visc_from_state = true;
    bool have_state = false;
    // launch geometry
    int ntx = 0, nty = 0, nzc = 0, zc = 0, nblocks = 0;
    Params base{};
    // profiling
    bool sync_check = false;  // CNS_SYNC_CHECK=1: synchronise and check after every launch
    // CNS_DIL_TILES=wide: the dilatation kernel always on its own 32 x 16
    // tile grid; =rhs: always on the RHS kernel's tiles; unset: by size
    int dil_tiles = 0;
    bool prof = false;
    std::vector<hipEvent_t> ev;
    int prof_n = 0;
    uint64_t dev_bytes = 0;
};

#define HIPCHK(s, expr)                                                                                   \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) {                                                                           \
            (s)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                                 \
            return CNS_ERR_HIP;                                                                           \
        }                                                                                                 \
    } while (0)

// After a launch: report launch errors; with sync_check also wait for the
// kernel and report its execution status, naming the kernel.
static int launched(cns_solver *s, const char *what) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && s->sync_check) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        s->err = std::string(what) + ": " + hipGetErrorString(e);
        return CNS_ERR_HIP;
    }
    return CNS_OK;
}
#define LAUNCHED(s, what)                   \
    do {                                    \
        int rc_ = launched((s), (what));    \
        if (rc_) return rc_;                \
    } while (0)

static int fail(cns_solver *s, int code, const std::string &msg) {
    if (s)
        s->err = msg;
    else
        g_err = msg;
    return code;
}

// Device blocks that a solver or batch releases stay cached per device and
// byte size, and the next allocation of the same size on that device takes
// one: hipFree waits for the whole device, and a serving loop closes and
// creates solvers of a few recurring sizes between its ticks.  At most
// cache_cap() bytes stay cached per device (CNS_CACHE_BYTES, default 16 GiB),
// the oldest blocks leave first; cns_release_cache frees them.  Blocks are
// released only after their work has finished (cns_destroy and
// cns_batch_destroy synchronise their streams first).
namespace {
struct CachedBlock {
    int dev;
    size_t bytes;
    void *p;
};
std::mutex g_cache_mu;
std::vector<CachedBlock> g_cache;                  // oldest first
std::unordered_map<void *, size_t> g_block_bytes;  // live blocks from dev_alloc
std::map<int, size_t> g_cached_bytes;              // per device
}  // namespace

static size_t cache_cap() {
    static const size_t cap = [] {
        const char *e = std::getenv("CNS_CACHE_BYTES");
        return e ? (size_t)std::strtoull(e, nullptr, 10) : ((size_t)16 << 30);
    }();
    return cap;
}

// frees cached blocks of device dev (the current device) until at most keep
// bytes stay; call with g_cache_mu held
static void cache_shrink(int dev, size_t keep) {
    for (size_t i = 0; i < g_cache.size() && g_cached_bytes[dev] > keep;) {
        if (g_cache[i].dev == dev) {
            (void)hipFree(g_cache[i].p);
            g_cached_bytes[dev] -= g_cache[i].bytes;
            g_cache.erase(g_cache.begin() + (long)i);
        } else {
            i++;
        }
    }
}

// allocation on the current device dev, from the cache when a block of the
// same size is there
static hipError_t dev_alloc(int dev, void **p, size_t bytes) {
    {
        std::lock_guard<std::mutex> lock(g_cache_mu);
        for (size_t i = g_cache.size(); i-- > 0;) {
            if (g_cache[i].dev == dev && g_cache[i].bytes == bytes) {
                *p = g_cache[i].p;
                g_cache.erase(g_cache.begin() + (long)i);
                g_cached_bytes[dev] -= bytes;
                g_block_bytes[*p] = bytes;
                return hipSuccess;
            }
        }
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) {
        std::lock_guard<std::mutex> lock(g_cache_mu);
        g_block_bytes[*p] = bytes;
    }
    return e;
}

// release of a block of the current device dev whose work has finished
static void dev_free(int dev, void *p) {
    if (!p) return;
    std::lock_guard<std::mutex> lock(g_cache_mu);
    auto it = g_block_bytes.find(p);
    if (it == g_block_bytes.end()) {
        (void)hipFree(p);
        return;
    }
    const size_t b = it->second;
    g_block_bytes.erase(it);
    if (b > cache_cap()) {
        (void)hipFree(p);
        return;
    }
    g_cache.push_back(CachedBlock{dev, b, p});
    g_cached_bytes[dev] += b;
    cache_shrink(dev, cache_cap());
}

extern "C" int cns_release_cache(int device) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    std::lock_guard<std::mutex> lock(g_cache_mu);
    std::vector<int> devs;
    for (const auto &kv : g_cached_bytes)
        if ((device < 0 || kv.first == device) && kv.second > 0) devs.push_back(kv.first);
    for (int d : devs) {
        if (hipSetDevice(d) != hipSuccess) continue;
        cache_shrink(d, 0);
    }
    (void)hipSetDevice(cur);
    return CNS_OK;
}

static int dalloc(cns_solver *s, double **p, size_t count) {
    HIPCHK(s, dev_alloc(s->dev, (void **)p, count * sizeof(double)));
    s->dev_bytes += count * sizeof(double);
    return CNS_OK;
}

static int balloc(cns_solver *s, void **p, size_t bytes) {
    HIPCHK(s, dev_alloc(s->dev, p, bytes));
    s->dev_bytes += bytes;
    return CNS_OK;
}

// grid and metrics: src/init.cpp:32-92 (initGrid), 258-277 (derivGrid),
// src/cuda_utils.cu:60-139 (inverse spacings, cell widths, stretched
// second-derivative weights)
static void build_grid(cns_solver *s) {
    const cns_config &c = s->cfg;
    const int mx = c.nx, my = c.ny, mz = c.nz, S = s->S, SV = s->SV;
    s->x.assign(mx, 0.0);
    s->xp.assign(mx, 0.0);
    s->xpp.assign(mx, 0.0);
    s->dxv.assign(mx, 0.0);
    s->y.assign(my, 0.0);
    s->z.assign(mz, 0.0);
    s->dx = c.len_x * (1.0) / (mx);
    // a boundary layer clusters x at the wall x = 0 only: the lower half of
    // the tanh profile (src/init.cpp:41-50)
    const int denom = c.boundary_layer ? 2 * mx : mx;
    const double denom2 = c.boundary_layer ? 1.0 : 2.0;
    std::vector<double> xn(mx + 1);
    for (int i = 0; i < mx + 1; i++) xn[i] = tanh(c.stretch * ((i * 1.0) / denom - 0.5)) / tanh(c.stretch * 0.5);
    for (int i = 0; i < mx; i++) s->x[i] = c.len_x * (1.0 + (xn[i] + xn[i + 1]) / 2.0) / denom2;
    // derivatives of x(xi) with odd/2Lx-odd mirrored ghosts, advective order
    std::vector<double> fb(mx + 2 * S);
    for (int i = S; i < mx + S; i++) fb[i] = s->x[i - S];
    for (int i = 0; i < S; i++) {
        fb[i] = -fb[2 * S - i - 1];
        fb[mx + S + i] = 2 * c.len_x - fb[mx + S - i - 1];
    }
    for (int i = 0; i < mx; i++) {
        double df = 0.0, d2f = s->cS[S] * fb[i + S] / s->dx / s->dx;
        for (int it = 0; it < S; it++) {
            df += s->cF[it] * (fb[i + it] - fb[i + S * 2 - it]) / s->dx;
            d2f += s->cS[it] * (fb[i + it] + fb[i + S * 2 - it]) / s->dx / s->dx;
        }
        s->xp[i] = 1.0 / df;
        s->xpp[i] = d2f;
    }
    if (!c.nonuniform_x) {
        for (int i = 0; i < mx; i++) s->x[i] = c.len_x * (0.5 + i * 1.0) / (mx);
        s->dx = s->x[1] - s->x[0];
    }
    for (int j = 0; j < my; j++) s->y[j] = c.len_y * (0.5 + j * 1.0) / (my);
    for (int k = 0; k < mz; k++) s->z[k] = c.len_z * (0.5 + k * 1.0) / (mz);
    s->d_dx = 1.0 / (s->dx);
    s->d_dy = 1.0 / (s->y[1] - s->y[0]);
    s->d_dz = 1.0 / (s->z[1] - s->z[0]);
    s->d_d2x = s->d_dx * s->d_dx;
    s->d_d2y = s->d_dy * s->d_dy;
    s->d_d2z = s->d_dz * s->d_dz;
    s->dxv[0] = (s->x[1] + s->x[0]) / 2.0;
    for (int i = 1; i < mx - 1; i++) s->dxv[i] = (s->x[i + 1] - s->x[i - 1]) / 2.0;
    s->dxv[mx - 1] = c.len_x - (s->x[mx - 1] + s->x[mx - 2]) / 2.0;
    // f_xx = xi_x^2 f_xixi + xi_xx f_xi folded into (2SV+1) x mx weights
    s->vsx.assign((size_t)mx * (2 * SV + 1), 0.0);
    const double *xp = s->xp.data(), *xpp = s->xpp.data();
    for (int it = 0; it < SV; it++)
        for (int i = 0; i < mx; i++)
            s->vsx[i + it * mx] = (s->cVS[it] * (xp[i] * xp[i]) * s->d_d2x -
                                   s->cVF[it] * xpp[i] * (xp[i] * xp[i] * xp[i]) * s->d_dx);
    for (int i = 0; i < mx; i++) s->vsx[i + SV * mx] = s->cVS[SV] * (xp[i] * xp[i]) * s->d_d2x;
    for (int it = SV + 1; it < 2 * SV + 1; it++)
        for (int i = 0; i < mx; i++)
            s->vsx[i + it * mx] = (s->cVS[2 * SV - it] * (xp[i] * xp[i]) * s->d_d2x +
                                   s->cVF[2 * SV - it] * xpp[i] * (xp[i] * xp[i] * xp[i]) * s->d_dx);
}

// Workgroup grid of the stage kernels: x-y tiles (tile_x() x 16 columns)
// times z chunks of zc planes.  Default chunk length (measured on 256^3 and
// 128^3 Taylor-Green cases): the FP64 kernel (one workgroup per CU) prefers
// long marches, about one workgroup per CU, chunks of 8..256 planes, and so
// does the FP32 kernel with mu = T/Re (one 512-lane workgroup per CU); small
// grids want chunks of 8.  cns_tune measures instead.
static void set_pencil(cns_solver *s, int zc) {
    const cns_config &c = s->cfg;
    zc = std::max(1, std::min(zc, c.nz));
    if (c.boundary_layer) {
        // every chunk at least S planes long: no window start of the RHS
        // kernel reaches past a z end (the kernels extrapolate only planes
        // they march over)
        zc = std::max(zc, s->S);
        while (c.nz % zc != 0 && c.nz % zc < s->S) zc++;
    }
    s->zc = zc;
    s->nzc = (c.nz + s->zc - 1) / s->zc;
    s->nblocks = s->ntx * s->nty * s->nzc;
    s->base.zc = s->zc;
    s->base.nzc = s->nzc;
}

static void launch_geometry(cns_solver *s) {
    const cns_config &c = s->cfg;
    const bool fp32 = c.precision == CNS_FP32, unitv = c.visc_exp == 1.0;
    const int tx = tile_x_of(fp32, unitv);
    s->ntx = (c.nx + tx - 1) / tx;
    s->nty = (c.ny + TILE_Y - 1) / TILE_Y;
    const long cols = (long)s->ntx * s->nty;
    // workgroups to aim at: one round, i.e. one 512-lane workgroup per CU at
    // two waves per SIMD, or two rounds of shorter marches for FP32 with a
    // general viscosity law (measured faster there)
    const long target = (fp32 && !unitv) ? 512 : 256;
    long zc = ((long)c.nz * cols + target - 1) / target;
    zc = std::max<long>(zc, 8);
    zc = std::min<long>(zc, 256);
    // CNS_PENCIL=<planes>: z-chunk length override
    if (const char *pz = std::getenv("CNS_PENCIL")) {
        const int v = std::atoi(pz);
        if (v >= 1) zc = v;
    }
    set_pencil(s, (int)zc);
}

static void fill_params(cns_solver *s) {
    const cns_config &c = s->cfg;
    Params &P = s->base;
    std::memset(&P, 0, sizeof(P));
    P.mx = c.nx;
    P.my = c.ny;
    P.mz = c.nz;
    P.zc = s->zc;
    P.ntx = s->ntx;
    P.nty = s->nty;
    P.nzc = s->nzc;
    P.tile_x = tile_x_of(s->fp32, c.visc_exp == 1.0);
    P.forcing = c.forcing_z;
    P.sponge = c.sponge;
    P.fix_y_work = c.fix_y_work;
    P.unit_visc = (c.visc_exp == 1.0);
    P.xcd_remap = 1;
    P.d_dx = s->d_dx;
    P.d_dy = s->d_dy;
    P.d_dz = s->d_dz;
    P.d_d2x = s->d_d2x;
    P.d_d2y = s->d_d2y;
    P.d_d2z = s->d_d2z;
    P.Rgas = s->Rgas;
    P.gam = c.gamma;
    P.cvInv = (c.gamma - 1.0) / s->Rgas;
    P.Re = c.reynolds;
    P.Pr = c.prandtl;
    P.Ec = s->Ec;
    P.viscexp = c.visc_exp;
    P.invPrEc = 1.0 / (c.prandtl * s->Ec);
    P.invRe = 1.0 / c.reynolds;
    P.cpR = s->Rgas * c.gamma / (c.gamma - 1.0);
    P.twall_bot = c.twall_bot;
    P.twall_top = c.twall_top;
    P.blayer = c.boundary_layer ? 1 : 0;
    P.perturbed = (c.boundary_layer && c.wall_blowing) ? 1 : 0;
    P.blow_om1 = c.reynolds * 121.e-6;
    P.blow_lam = c.len_y / (2.0 * M_PI);
    for (int i = 0; i < 4; i++) {
        P.cF[i] = s->cF[i];
        P.cVF[i] = s->cVF[i];
    }
    for (int i = 0; i < 5; i++) {
        P.cS[i] = s->cS[i];
        P.cVS[i] = s->cVS[i];
    }
    P.xp = s->d_xp;
    P.vsx = s->d_vsx;
    P.dxv = s->d_dxv;
    P.spx = s->d_spx;
    P.spz = s->d_spz;
    P.ref = s->d_ref;
    P.dt = &s->d_sc->dt;
    P.dpdz = &s->d_sc->dpdz;
    P.time = &s->d_sc->time;
    P.visc_max_out = &s->d_sc->visc_max;
    P.comp = s->comp ? 1 : 0;
}

extern "C" int cns_config_init(cns_config *c) {
    if (!c) return fail(nullptr, CNS_ERR_ARG, "cns_config_init: null config");
    std::memset(c, 0, sizeof(*c));
    c->nx = c->ny = c->nz = 64;
    c->stencil_size = 2;
    c->stencil_visc = 2;
    c->periodic_x = 1;
    c->nonuniform_x = 0;
    c->forcing_z = 0;
    c->sponge = 0;
    c->scheme = CNS_RK3_LOWSTORAGE;
    c->check_cfl = 10;
    c->check_bulk = 10;
    c->fix_y_work = 0;
    c->precision = CNS_FP64;
    c->compensated = 1;
    c->device = 0;
    c->len_x = c->len_y = c->len_z = 2.0 * M_PI;
    c->reynolds = 1600.0;
    c->prandtl = 0.71;
    c->mach = 0.1;
    c->gamma = 1.4;
    c->visc_exp = 1.0;
    c->cfl = 0.5;
    c->stretch = 1.8;
    c->twall_bot = 1.0;
    c->twall_top = 1.0;
    c->dpdz0 = 0.00372;
    c->boundary_layer = 0;
    c->wall_blowing = 0;
    return CNS_OK;
}

extern "C" const char *cns_last_error(const cns_solver *s) { return s ? s->err.c_str() : g_err.c_str(); }

static int validate(const cns_config *c, std::string &msg) {
    if (c->stencil_size < 1 || c->stencil_size > 4) {
        msg = "stencil_size must be 1..4 (orders 2, 4, 6, 8)";
        return CNS_ERR_ARG;
    }
    if (c->stencil_visc < 1 || c->stencil_visc > c->stencil_size) {
        msg = "stencil_visc must be 1..stencil_size";
        return CNS_ERR_ARG;
    }
    const int need = 2 * c->stencil_size;
    if (c->nx < need || c->ny < need || c->nz < need) {
        msg = "every grid dimension must be at least 2*stencil_size";
        return CNS_ERR_ARG;
    }
    if ((long)c->nx * c->ny * c->nz > (1L << 31) - 1) {
        msg = "grid larger than 2^31-1 points";
        return CNS_ERR_ARG;
    }
    if ((long)c->nx * c->ny > (1L << 29)) {
        // the kernels address a plane with 32-bit byte offsets
        msg = "x-y plane larger than 2^29 points";
        return CNS_ERR_ARG;
    }
    if (c->scheme < 0 || c->scheme > 2) {
        msg = "scheme must be CNS_RK3_LOWSTORAGE, CNS_RK3_KUTTA or CNS_RK4";
        return CNS_ERR_ARG;
    }
    if (c->check_cfl < 1 || c->check_bulk < 1) {
        msg = "check_cfl and check_bulk must be >= 1";
        return CNS_ERR_ARG;
    }
    if (!(c->len_x > 0 && c->len_y > 0 && c->len_z > 0 && c->reynolds > 0 && c->prandtl > 0 && c->mach > 0 && c->gamma > 1 &&
          c->cfl > 0)) {
        msg = "lengths, Re, Pr, Ma, cfl must be positive and gam > 1";
        return CNS_ERR_ARG;
    }
    if (c->precision != CNS_FP64 && c->precision != CNS_FP32) {
        msg = "precision must be CNS_FP64 or CNS_FP32";
        return CNS_ERR_ARG;
    }
    if (c->boundary_layer) {
        if (c->periodic_x) {
            msg = "boundary_layer needs periodic_x = 0";
            return CNS_ERR_ARG;
        }
        if (c->precision != CNS_FP64) {
            msg = "boundary_layer runs with FP64 storage only";
            return CNS_ERR_ARG;
        }
        if (c->nx % 16 != 0) {
            // the top x ghosts are formed from the last 16-wide tile's nodes
            msg = "boundary_layer needs nx to be a multiple of 16";
            return CNS_ERR_ARG;
        }
        if (c->nz <= 2 * c->stencil_size) {
            msg = "boundary_layer needs nz > 2*stencil_size";
            return CNS_ERR_ARG;
        }
    }
    return CNS_OK;
}

extern "C" int cns_destroy(cns_solver *s) {
    if (!s) return CNS_OK;
    if (s->dev >= 0) (void)hipSetDevice(s->dev);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    const int dev = s->dev;
    auto fr = [dev](double *p) { dev_free(dev, p); };
    fr(s->d_xp);
    fr(s->d_x);
    fr(s->d_y);
    fr(s->d_z);
    fr(s->d_vsx);
    fr(s->d_dxv);
    fr(s->d_spx);
    fr(s->d_spz);
    fr(s->d_ref);
    dev_free(dev, s->d_sc);
    dev_free(dev, s->d_ctl);
    fr(s->d_part);
    auto fv = [dev](void *p) { dev_free(dev, p); };
    fv(s->d_dil);
    for (int a = 0; a < 4; a++) {
        fv(s->sets[a]);
        fv(s->lo[a]);
    }
    for (int a = 0; a < 2; a++)
        fv(s->rbuf[a]);
    fr(s->d_par1);
    fr(s->d_par2);
    fr(s->d_time);
    fr(s->d_prof);
    fr(s->d_stats);
    for (auto &e : s->ev) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return CNS_OK;
}

extern "C" int cns_create(const cns_config *cfg, cns_solver **out) {
    if (!cfg || !out) return fail(nullptr, CNS_ERR_ARG, "cns_create: null argument");
    *out = nullptr;
    std::string msg;
    int rc = validate(cfg, msg);
    if (rc) return fail(nullptr, rc, msg);
    int ndev = 0;
    const hipError_t de = hipGetDeviceCount(&ndev);
    if (de != hipSuccess || ndev <= 0)
        return fail(nullptr, CNS_ERR_HIP, std::string("cns_create: no HIP device available (hipGetDeviceCount: ") +
                                              hipGetErrorString(de) + ", count " + std::to_string(ndev) + ")");
    if (cfg->device < 0 || cfg->device >= ndev) return fail(nullptr, CNS_ERR_ARG, "cns_create: bad device ordinal");
    cns_solver *s = new cns_solver();
    s->cfg = *cfg;
    s->dev = cfg->device;
    {
        const char *sc = std::getenv("CNS_SYNC_CHECK");
        s->sync_check = sc && sc[0] == '1';
        const char *dt = std::getenv("CNS_DIL_TILES");
        s->dil_tiles = dt ? (std::strcmp(dt, "wide") == 0 ? 1 : (std::strcmp(dt, "rhs") == 0 ? 2 : 0)) : 0;
    }
    s->S = cfg->stencil_size;
    s->SV = cfg->stencil_visc;
    s->fp32 = cfg->precision == CNS_FP32;
    s->comp = s->fp32 && cfg->compensated != 0;
    s->esz = s->fp32 ? sizeof(float) : sizeof(double);
    s->n = (long)cfg->nx * cfg->ny * cfg->nz;
    fd_weights(s->S, s->cF, s->cS);
    fd_weights(s->SV, s->cVF, s->cVS);
    s->Ec = ((cfg->gamma - 1.0) * cfg->mach * cfg->mach);
    s->Rgas = (1.0 / (cfg->gamma * cfg->mach * cfg->mach));
    build_grid(s);
    launch_geometry(s);
    auto bail = [&](int code) {
        g_err = s->err;
        cns_destroy(s);
        return code;
    };
    if (hipSetDevice(s->dev) != hipSuccess) {
        s->err = "hipSetDevice failed";
        return bail(CNS_ERR_HIP);
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        s->err = "hipStreamCreate failed";
        return bail(CNS_ERR_HIP);
    }
    const int mx = cfg->nx, mz = cfg->nz;
    const long n = s->n;
    s->nsets = (cfg->scheme == CNS_RK3_LOWSTORAGE) ? 2 : (cfg->scheme == CNS_RK3_KUTTA ? 3 : 4);
    s->nrhs = (cfg->scheme == CNS_RK4) ? 1 : 2;
    if ((rc = dalloc(s, &s->d_x, mx)) || (rc = dalloc(s, &s->d_y, cfg->ny)) || (rc = dalloc(s, &s->d_z, mz)))
        return bail(rc);
    if ((rc = dalloc(s, &s->d_xp, mx)) || (rc = dalloc(s, &s->d_vsx, s->vsx.size())) ||
        (rc = dalloc(s, &s->d_dxv, mx)) || (rc = dalloc(s, &s->d_spx, mx)) || (rc = dalloc(s, &s->d_spz, mz)) ||
        (rc = dalloc(s, &s->d_part, (size_t)NPART * RED_BLOCKS)) || (rc = balloc(s, &s->d_dil, n * s->esz)))
        return bail(rc);
    if ((rc = dalloc(s, &s->d_ref, (size_t)5 * mx * mz))) return bail(rc);
    for (int a = 0; a < s->nsets; a++) {
        if ((rc = balloc(s, &s->sets[a], 5 * (size_t)n * s->esz))) return bail(rc);
        // low parts only where a running sum lives: both SMR sets (they
        // alternate), the Kutta step state (set 0), the RK4 state and
        // accumulator (sets 0 and 2); intermediate stage inputs need none
        const bool need = cfg->scheme == CNS_RK3_LOWSTORAGE || a == 0 || (cfg->scheme == CNS_RK4 && a == 2);
        if (s->comp && need && (rc = balloc(s, &s->lo[a], 5 * (size_t)n * s->esz))) return bail(rc);
    }
    for (int a = 0; a < s->nrhs; a++)
        if ((rc = balloc(s, &s->rbuf[a], 5 * (size_t)n * s->esz))) return bail(rc);
    if (dev_alloc(s->dev, (void **)&s->d_ctl, sizeof(StepCtl)) != hipSuccess) {
        s->err = "hipMalloc(step control) failed";
        return bail(CNS_ERR_NOMEM);
    }
    if (dev_alloc(s->dev, (void **)&s->d_sc, sizeof(Scalars)) != hipSuccess) {
        s->err = "hipMalloc(scalars) failed";
        return bail(CNS_ERR_NOMEM);
    }
    Scalars h{};
    h.dt = 0.0;
    h.dpdz = cfg->forcing_z ? cfg->dpdz0 : 0.0;
    h.time = 0.0;
    if (hipMemcpy(s->d_sc, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_xp, s->xp.data(), mx * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_x, s->x.data(), mx * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_y, s->y.data(), cfg->ny * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_z, s->z.data(), mz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_vsx, s->vsx.data(), s->vsx.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_dxv, s->dxv.data(), mx * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        s->err = "initial host-to-device copies failed";
        return bail(CNS_ERR_HIP);
    }
    std::vector<double> zeros((size_t)mx * mz > (size_t)mz ? (size_t)mx * mz : (size_t)mz, 0.0);
    if (hipMemcpy(s->d_spx, zeros.data(), mx * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->d_spz, zeros.data(), mz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
        s->err = "sponge initialisation failed";
        return bail(CNS_ERR_HIP);
    }
    for (int q = 0; q < 5; q++)
        if (hipMemcpy(s->d_ref + (size_t)q * mx * mz, zeros.data(), (size_t)mx * mz * sizeof(double),
                      hipMemcpyHostToDevice) != hipSuccess) {
            s->err = "sponge initialisation failed";
            return bail(CNS_ERR_HIP);
        }
    fill_params(s);
    *out = s;
    return CNS_OK;
}

extern "C" int cns_get_grid(const cns_solver *s, double *x, double *y, double *z, double *xp, double *xpp,
                            double *dxv) {
    if (!s) return fail(nullptr, CNS_ERR_ARG, "null solver");
    const int mx = s->cfg.nx;
    if (x) std::memcpy(x, s->x.data(), mx * sizeof(double));
    if (xp) std::memcpy(xp, s->xp.data(), mx * sizeof(double));
    if (xpp) std::memcpy(xpp, s->xpp.data(), mx * sizeof(double));
    if (dxv) std::memcpy(dxv, s->dxv.data(), mx * sizeof(double));
    if (y) std::memcpy(y, s->y.data(), s->cfg.ny * sizeof(double));
    if (z) std::memcpy(z, s->z.data(), s->cfg.nz * sizeof(double));
    return CNS_OK;
}

// src/init.cpp:126-148 (initCHIT): V0 = T0 = R0 = 1, P0 = T0*Rgas
extern "C" int cns_fill_taylor_green(const cns_solver *s, double *r, double *u, double *v, double *w, double *e) {
    if (!s || !r || !u || !v || !w || !e) return fail(nullptr, CNS_ERR_ARG, "cns_fill_taylor_green: null argument");
    const cns_config &c = s->cfg;
    const double V0 = 1.0, T0 = 1.0, P0 = T0 * s->Rgas, R0 = 1.0;
    const int mx = c.nx, my = c.ny, mz = c.nz;
    host_parallel((long)my * mz, [&](long a, long b) {
        for (long jk = a; jk < b; jk++) {
            const int j = (int)(jk % my), k = (int)(jk / my);
            const double fz = 2 * M_PI * s->z[k] / c.len_z;
            const double fy = 2 * M_PI * s->y[j] / c.len_y;
            for (int i = 0; i < mx; i++) {
                const double fx = 2 * M_PI * s->x[i] / c.len_x;
                const long g = i + (long)mx * (j + (long)my * k);
                u[g] = V0 * sin(fx / 1.0) * cos(fy / 1.0) * cos(fz / 1.0);
                v[g] = -V0 * cos(fx / 1.0) * sin(fy / 1.0) * cos(fz / 1.0);
                w[g] = 0.0;
                const double press = P0 + 1.0 / 16.0 * R0 * V0 * V0 * (cos(2.0 * fx / 1.0) + cos(2.0 * fy / 1.0)) *
                                              (cos(2.0 * fz / 1.0) + 2.0);
                r[g] = press / s->Rgas / T0;
                e[g] = press / (c.gamma - 1.0) + 0.5 * r[g] * (pow(u[g], 2) + pow(v[g], 2) + pow(w[g], 2));
            }
        }
    });
    return CNS_OK;
}

// Isothermal channel: rho = 1, T = 1, p = Rgas; streamwise (z) velocity
// 1.5*ubulk*(1-(x/h-1)^2) with h = Lx/2; every component perturbed by
// amp*ubulk*(2U-1), U from SplitMix64(seed) drawn in storage order, three
// draws (u, v, w) per point.
extern "C" int cns_fill_channel(const cns_solver *s, uint64_t seed, double ubulk, double amp, double *r, double *u,
                                double *v, double *w, double *e) {
    if (!s || !r || !u || !v || !w || !e) return fail(nullptr, CNS_ERR_ARG, "cns_fill_channel: null argument");
    const cns_config &c = s->cfg;
    const double T0 = 1.0, P0 = T0 * s->Rgas, R0 = 1.0, hh = c.len_x / 2.0;
    const int mx = c.nx, my = c.ny, mz = c.nz;
    // SplitMix64's state after m draws is seed + m * 0x9E3779B97F4A7C15, so
    // point g's three draws start from seed + 3 g times that increment: the
    // same numbers as one sequential stream, drawn in parallel
    host_parallel((long)mx * my * mz, [&](long a, long b) {
        SplitMix64 rng{seed + (uint64_t)(3 * a) * 0x9E3779B97F4A7C15ull};
        for (long g = a; g < b; g++) {
            const int i = (int)(g % mx);
            const double xi = s->x[i] / hh - 1.0;
            const double wm = 1.5 * ubulk * (1.0 - xi * xi);
            const double a1 = rng.uniform(), a2 = rng.uniform(), a3 = rng.uniform();
            u[g] = amp * ubulk * (2.0 * a1 - 1.0);
            v[g] = amp * ubulk * (2.0 * a2 - 1.0);
            w[g] = wm + amp * ubulk * (2.0 * a3 - 1.0);
            r[g] = R0;
            e[g] = P0 / (c.gamma - 1.0) + 0.5 * r[g] * (u[g] * u[g] + v[g] * v[g] + w[g] * w[g]);
        }
    });
    return CNS_OK;
}

template <typename R>
static R *as(void *p) {
    return static_cast<R *>(p);
}

// Pointers a stage does not use still point at valid device memory (the
// stage input), so no kernel argument is ever a null device address.
template <typename R>
static void fill_unused(StagePtrs<R> &A) {
    R *in = const_cast<R *>(A.qin);
    if (!A.s1) A.s1 = in;
    if (!A.ra) A.ra = in;
    if (!A.rb) A.rb = in;
    if (!A.s2) A.s2 = in;
    if (!A.qout) A.qout = in;
    if (!A.rout) A.rout = in;
    if (!A.accout) A.accout = in;
    if (!A.s1lo) A.s1lo = in;
    if (!A.s2lo) A.s2lo = in;
    if (!A.qoutlo) A.qoutlo = in;
    if (!A.accoutlo) A.accoutlo = in;
}

template <typename R>
static StagePtrs<R> ptrs_in(cns_solver *s, int set) {
    StagePtrs<R> A;
    std::memset(&A, 0, sizeof(A));
    A.qin = as<R>(s->sets[set]);
    A.fs = s->n;
    A.dil = as<R>(s->d_dil);
    fill_unused(A);
    return A;
}

// FP32 storage: the reference state of the offsets is the mean density and
// total energy of the state given to cns_set_state.
static void set_offsets(cns_solver *s, double rho0, double E0) {
    const double g = s->cfg.gamma;
    s->rho0 = rho0;
    s->E0 = E0;
    Params &P = s->base;
    P.rho0 = rho0;
    P.E0 = E0;
    const double p0 = (g - 1.0) * E0;
    P.T0 = p0 / (rho0 * s->Rgas);
    P.h0 = (E0 + p0) / rho0;
    P.dtw_bot = s->cfg.twall_bot - P.T0;
    P.dtw_top = s->cfg.twall_top - P.T0;
    P.ch0 = P.cpR * P.T0 - P.h0;
}

extern "C" int cns_set_state(cns_solver *s, const double *r, const double *u, const double *v, const double *w,
                             const double *e) {
    if (!s || !r || !u || !v || !w || !e) return fail(s, CNS_ERR_ARG, "cns_set_state: null argument");
    if (s->in_batch) return fail(s, CNS_ERR_STATE, "cns_set_state: the solver is advanced by a batch");
    HIPCHK(s, hipSetDevice(s->dev));
    const long n = s->n;
    s->cur = 0;
    if (!s->fp32) {
        double *const set = as<double>(s->sets[0]);
        const size_t b = n * sizeof(double);
        const double *src[5] = {r, u, v, w, e};
        for (int q = 0; q < 5; q++)
            HIPCHK(s, hipMemcpyAsync(set + q * n, src[q], b, hipMemcpyHostToDevice, s->stream));
        hipLaunchKernelGGL(k_to_conserved<double>, dim3(1024), dim3(256), 0, s->stream, n, set, set + n, set + 2 * n,
                           set + 3 * n);
        LAUNCHED(s, "k_to_conserved<double>");
    } else {
        // conserved variables in FP64 as the FP64 path forms them (rho u = u rho),
        // minus the offsets, split into hi = fl(x) and lo = fl(x - hi)
        double sr = 0.0, se = 0.0;
        for (long g = 0; g < n; g++) {
            sr += r[g];
            se += e[g];
        }
        set_offsets(s, sr / n, se / n);
        std::vector<float> hi(5 * (size_t)n), lo(5 * (size_t)n);
        host_parallel(n, [&](long a, long b) {
            for (long g = a; g < b; g++) {
                const double x[5] = {r[g] - s->rho0, u[g] * r[g], v[g] * r[g], w[g] * r[g], e[g] - s->E0};
                for (int q = 0; q < 5; q++) {
                    const float h = (float)x[q];
                    hi[q * n + g] = h;
                    lo[q * n + g] = (float)(x[q] - (double)h);
                }
            }
        });
        HIPCHK(s, hipMemcpy(s->sets[0], hi.data(), hi.size() * sizeof(float), hipMemcpyHostToDevice));
        if (s->comp) HIPCHK(s, hipMemcpy(s->lo[0], lo.data(), lo.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    s->bsteps = 0;
    // time restarts at zero (initDevice, src/cuda_utils.cu:386)
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(1), 0, s->stream, &s->d_sc->time, 1L, 0.0);
    LAUNCHED(s, "k_fill");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->visc_from_state = true;
    s->have_state = true;
    return CNS_OK;
}

// initial condition evaluated on the device (kind 0 Taylor-Green, 1 channel)
static int init_on_device(cns_solver *s, int kind, uint64_t seed, double ubulk, double amp) {
    if (s->in_batch) return fail(s, CNS_ERR_STATE, "cns_init: the solver is advanced by a batch");
    HIPCHK(s, hipSetDevice(s->dev));
    const cns_config &c = s->cfg;
    InitArgs a{};
    a.kind = kind;
    a.mx = c.nx;
    a.my = c.ny;
    a.mz = c.nz;
    a.x = s->d_x;
    a.y = s->d_y;
    a.z = s->d_z;
    a.lx = c.len_x;
    a.ly = c.len_y;
    a.lz = c.len_z;
    a.Rgas = s->Rgas;
    a.gam = c.gamma;
    a.ubulk = ubulk;
    a.amp = amp;
    a.seed = seed;
    const long n = s->n;
    s->cur = 0;
    if (!s->fp32) {
        hipLaunchKernelGGL(k_init_store<double>, dim3(1024), dim3(256), 0, s->stream, a, n, as<double>(s->sets[0]),
                           (double *)nullptr, 0.0, 0.0);
        LAUNCHED(s, "k_init_store<double>");
    } else {
        // reference state: the mean density and total energy, summed in a fixed order
        hipLaunchKernelGGL(k_init_sums, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s->stream, a, n, s->d_part);
        LAUNCHED(s, "k_init_sums");
        std::vector<double> part(2 * RED_BLOCKS);
        HIPCHK(s, hipMemcpyAsync(part.data(), s->d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                                 s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        double sr = 0.0, se = 0.0;
        for (int b = 0; b < RED_BLOCKS; b++) {
            sr += part[b];
            se += part[RED_BLOCKS + b];
        }
        set_offsets(s, sr / n, se / n);
        hipLaunchKernelGGL(k_init_store<float>, dim3(1024), dim3(256), 0, s->stream, a, n, as<float>(s->sets[0]),
                           s->comp ? as<float>(s->lo[0]) : (float *)nullptr, s->rho0, s->E0);
        LAUNCHED(s, "k_init_store<float>");
    }
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(1), 0, s->stream, &s->d_sc->time, 1L, 0.0);
    LAUNCHED(s, "k_fill");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->bsteps = 0;
    s->visc_from_state = true;
    s->have_state = true;
    return CNS_OK;
}

extern "C" int cns_init_taylor_green(cns_solver *s) {
    if (!s) return fail(nullptr, CNS_ERR_ARG, "null solver");
    return init_on_device(s, 0, 0, 0.0, 0.0);
}

extern "C" int cns_init_channel(cns_solver *s, uint64_t seed, double ubulk, double amp) {
    if (!s) return fail(nullptr, CNS_ERR_ARG, "null solver");
    return init_on_device(s, 1, seed, ubulk, amp);
}

extern "C" int cns_get_state(cns_solver *s, double *r, double *u, double *v, double *w, double *e) {
    if (!s || !r || !u || !v || !w || !e) return fail(s, CNS_ERR_ARG, "cns_get_state: null argument");
    if (!s->have_state) return fail(s, CNS_ERR_STATE, "cns_get_state: no state set");
    HIPCHK(s, hipSetDevice(s->dev));
    const long n = s->n;
    if (!s->fp32) {
        const size_t b = n * sizeof(double);
        double *const set = as<double>(s->sets[s->cur]);
        double *const q[5] = {set, set + n, set + 2 * n, set + 3 * n, set + 4 * n};
        // velocities into the first rhs buffer
        double *tu = as<double>(s->rbuf[0]), *tv = tu + n, *tw = tu + 2 * n;
        hipLaunchKernelGGL(k_to_velocity<double>, dim3(1024), dim3(256), 0, s->stream, n, q[0], q[1], q[2], q[3], tu,
                           tv, tw);
        LAUNCHED(s, "k_to_velocity<double>");
        HIPCHK(s, hipMemcpyAsync(r, q[0], b, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipMemcpyAsync(u, tu, b, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipMemcpyAsync(v, tv, b, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipMemcpyAsync(w, tw, b, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipMemcpyAsync(e, q[4], b, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
        return CNS_OK;
    }
    std::vector<float> hi(5 * (size_t)n), lo(5 * (size_t)n, 0.0f);
    // on the solver's own stream: other streams' work (a batch) keeps running
    HIPCHK(s, hipMemcpyAsync(hi.data(), s->sets[s->cur], hi.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    if (s->comp)
        HIPCHK(s, hipMemcpyAsync(lo.data(), s->lo[s->cur], lo.size() * sizeof(float), hipMemcpyDeviceToHost,
                                 s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    host_parallel(n, [&](long a, long b) {
        for (long g = a; g < b; g++) {
            double x[5];
            for (int q = 0; q < 5; q++) x[q] = (double)hi[q * n + g] + (double)lo[q * n + g];
            r[g] = x[0] + s->rho0;
            u[g] = x[1] / r[g];
            v[g] = x[2] / r[g];
            w[g] = x[3] / r[g];
            e[g] = x[4] + s->E0;
        }
    });
    return CNS_OK;
}

// dilatation + fused RHS/update of one stage on the stage input qin, with
// the dilatation field in dil; with timing, the RHS launch is bracketed by
// events when profiling is on
template <typename R>
static int run_stage_on(cns_solver *s, const R *qin, R *dil, Params P, StagePtrs<R> A, bool timing) {
    const bool wallx = !s->cfg.periodic_x, nonuni = s->cfg.nonuniform_x != 0;
    A.qin = qin;
    A.fs = s->n;
    A.dil = dil;
    fill_unused(A);
    if constexpr (std::is_same<R, double>::value) {
        if (s->cfg.boundary_layer) {
            // boundary layer: both kernels on the RHS kernel's 16 x 16 tiles
            // and z chunks (the last x tile holds the nodes the top ghosts
            // extrapolate; every chunk is at least S planes long)
            HIPCHK(s, launch_dil_bl(s->SV, nonuni, s->nblocks, s->stream, P, A, dil));
            LAUNCHED(s, "k_dil");
            if (P.want_visc_max) {
                hipLaunchKernelGGL(k_reset_visc, dim3(1), dim3(1), 0, s->stream, s->d_sc);
                LAUNCHED(s, "k_reset_visc");
            }
            const bool prof = s->prof && timing;
            if (prof) {
                const size_t need = 2 * (size_t)(s->prof_n + 1);
                while (s->ev.size() < need) {
                    hipEvent_t e;
                    HIPCHK(s, hipEventCreate(&e));
                    s->ev.push_back(e);
                }
                HIPCHK(s, hipEventRecord(s->ev[2 * s->prof_n], s->stream));
            }
            HIPCHK(s, launch_rhs_bl(s->S, s->SV, nonuni, s->nblocks, s->stream, P, A));
            LAUNCHED(s, "k_rhs");
            if (prof) {
                HIPCHK(s, hipEventRecord(s->ev[2 * s->prof_n + 1], s->stream));
                s->prof_n++;
            }
            return CNS_OK;
        }
    }
    // the dilatation kernel needs few registers, so it always runs 512-lane
    // workgroups on 32 x 16 tiles (fewer halo points per point), with its own
    // tile grid and z chunks: about 1024 workgroups, chunks of 8..the RHS's
    // (small grids, where that leaves fewer than 512 workgroups, keep the
    // RHS kernel's tiles; CNS_DIL_TILES overrides the choice)
    Params Pd = P;
    Pd.tile_x = 32;
    Pd.ntx = (s->cfg.nx + 31) / 32;
    {
        const long dcols = (long)Pd.ntx * Pd.nty;
        const int dz = (int)std::max<long>(8, std::min<long>(s->zc, (long)s->cfg.nz * dcols / 1024));
        Pd.zc = std::min(dz, s->cfg.nz);
        Pd.nzc = (s->cfg.nz + Pd.zc - 1) / Pd.zc;
    }
    int dil_blocks = Pd.ntx * Pd.nty * Pd.nzc;
    if (s->dil_tiles == 2 || (dil_blocks < 512 && s->dil_tiles != 1)) {
        Pd = P;
        dil_blocks = s->nblocks;
    }
    HIPCHK(s, launch_dil<R>(s->SV, wallx, nonuni, dil_blocks, s->stream, Pd, A, dil));
    LAUNCHED(s, "k_dil");
    if (P.want_visc_max) {
        hipLaunchKernelGGL(k_reset_visc, dim3(1), dim3(1), 0, s->stream, s->d_sc);
        LAUNCHED(s, "k_reset_visc");
    }
    const bool prof = s->prof && timing;
    if (prof) {
        const size_t need = 2 * (size_t)(s->prof_n + 1);
        while (s->ev.size() < need) {
            hipEvent_t e;
            HIPCHK(s, hipEventCreate(&e));
            s->ev.push_back(e);
        }
        HIPCHK(s, hipEventRecord(s->ev[2 * s->prof_n], s->stream));
    }
    HIPCHK(s, launch_rhs<R>(s->S, s->SV, wallx, nonuni, s->nblocks, s->stream, P, A));
    LAUNCHED(s, "k_rhs");
    if (prof) {
        HIPCHK(s, hipEventRecord(s->ev[2 * s->prof_n + 1], s->stream));
        s->prof_n++;
    }
    return CNS_OK;
}

template <typename R>
static int run_stage(cns_solver *s, int in_set, Params P, StagePtrs<R> A) {
    return run_stage_on<R>(s, as<R>(s->sets[in_set]), as<R>(s->d_dil), P, A, true);
}

template <typename R>
static int rhs_t(cns_solver *s, double *const dst[5]) {
    Params P = s->base;
    P.write_q = 0;
    P.write_r = 1;
    P.want_visc_max = 1;
    StagePtrs<R> A;
    std::memset(&A, 0, sizeof(A));
    A.rout = as<R>(s->rbuf[0]);
    int rc = run_stage<R>(s, s->cur, P, A);
    if (rc) return rc;
    s->visc_from_state = false;
    const long n = s->n;
    if (sizeof(R) == sizeof(double)) {
        for (int q = 0; q < 5; q++)
            HIPCHK(s, hipMemcpyAsync(dst[q], as<double>(s->rbuf[0]) + q * n, n * sizeof(double),
                                     hipMemcpyDeviceToHost, s->stream));
        HIPCHK(s, hipStreamSynchronize(s->stream));
    } else {
        std::vector<R> h(5 * (size_t)n);
        HIPCHK(s, hipStreamSynchronize(s->stream));
        HIPCHK(s, hipMemcpy(h.data(), s->rbuf[0], h.size() * sizeof(R), hipMemcpyDeviceToHost));
        for (int q = 0; q < 5; q++)
            for (long g = 0; g < n; g++) dst[q][g] = (double)h[q * n + g];
    }
    return CNS_OK;
}

// Scratch of cns_rhs_noalloc: the dilatation field of one evaluation, in a
// file-scope device array (16 MiB on every GPU the library runs on, held
// until the process ends), so that the entry point allocates nothing.  Host
// code reaches it only through the per-device address hipGetSymbolAddress
// returns (the array's name on the host is a stand-in, not a device
// address).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
}  // namespace

// Device address of the scratch array on device dev (current), looked up once
// per device.
static int scratch_address(cns_solver *s, int dev, void **out) {
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(s, CNS_ERR_UNSUPPORTED, "cns_rhs_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(s, CNS_ERR_HIP, std::string("cns_rhs_noalloc: hipGetSymbolAddress(scratch): ") +
                                            hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return CNS_OK;
}

template <typename R>
static int rhs_noalloc_t(cns_solver *s, const void *q, void *rhs, void *dil) {
    Params P = s->base;
    P.write_q = 0;
    P.write_r = 1;
    P.want_visc_max = 0;
    StagePtrs<R> A;
    std::memset(&A, 0, sizeof(A));
    A.rout = static_cast<R *>(rhs);
    return run_stage_on<R>(s, static_cast<const R *>(q), static_cast<R *>(dil), P, A, false);
}

extern "C" int cns_rhs_noalloc(cns_solver *s, const void *q, void *rhs) {
    if (!s || !q || !rhs) return fail(s, CNS_ERR_ARG, "cns_rhs_noalloc: null argument");
    if (s->in_batch) return fail(s, CNS_ERR_STATE, "cns_rhs_noalloc: the solver is advanced by a batch");
    const size_t need = (size_t)s->n * (s->fp32 ? sizeof(float) : sizeof(double));
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(s, CNS_ERR_ARG, "cns_rhs_noalloc: " + std::to_string(s->n) + " points need " + std::to_string(need) +
                                        " bytes of scratch, above its maximum of " +
                                        std::to_string(NOALLOC_SCRATCH_BYTES) + " bytes");
    HIPCHK(s, hipSetDevice(s->dev));
    void *dil = nullptr;
    int rc = scratch_address(s, s->dev, &dil);
    if (rc) return rc;
    return s->fp32 ? rhs_noalloc_t<float>(s, q, rhs, dil) : rhs_noalloc_t<double>(s, q, rhs, dil);
}

extern "C" int cns_rhs(cns_solver *s, double *rr, double *ru, double *rv, double *rw, double *re) {
    if (!s || !rr || !ru || !rv || !rw || !re) return fail(s, CNS_ERR_ARG, "cns_rhs: null argument");
    if (s->in_batch) return fail(s, CNS_ERR_STATE, "cns_rhs: the solver is advanced by a batch");
    if (!s->have_state) return fail(s, CNS_ERR_STATE, "cns_rhs: no state set");
    HIPCHK(s, hipSetDevice(s->dev));
    double *const dst[5] = {rr, ru, rv, rw, re};
    return s->fp32 ? rhs_t<float>(s, dst) : rhs_t<double>(s, dst);
}

// Launch parameters of one stage update
// q_out = s1 + dt (w_ra ra + w_rb rb + w_r rhs).
static Params stage_params(const cns_solver *s, double w_ra, double w_rb, double w_r) {
    Params P = s->base;
    P.write_q = 1;
    P.w_ra = w_ra;
    P.w_rb = w_rb;
    P.w_r = w_r;
    P.use_ra = w_ra != 0.0;
    P.use_rb = w_rb != 0.0;
    return P;
}

static const double SMR_ALPHA[3] = {0., -17. / 60., -5. / 12.};
static const double SMR_BETA[3] = {8. / 15., 5. / 12., 3. / 4.};

static int n_stages(const cns_solver *s) { return s->cfg.scheme == CNS_RK4 ? 4 : 3; }

// set holding the state after a step that started from set cur (the
// low-storage RK3 alternates between its two sets)
static int next_cur(const cns_solver *s, int cur) { return s->cfg.scheme == CNS_RK3_LOWSTORAGE ? 1 - cur : cur; }

// Parameters and arrays of stage st of a step that starts from set cur;
// returns the set the stage reads with stencils.  last_visc: the last stage
// accumulates the viscous CFL limit for the next check.
template <typename R>
static int stage_setup(const cns_solver *s, int cur, int st, bool last_visc, Params &P, StagePtrs<R> &A) {
    const cns_config &c = s->cfg;
    std::memset(&A, 0, sizeof(A));
    int in = cur;
    if (c.scheme == CNS_RK3_LOWSTORAGE) {
        // q <- q + dt (alpha_s r_prev + beta_s r_new)  (sumLowStorageRK3)
        const int a = cur, b = 1 - cur;
        const int ins[3] = {a, b, a}, outs[3] = {b, a, b};
        in = ins[st];
        P = stage_params(s, st ? SMR_ALPHA[st] : 0, 0, SMR_BETA[st]);
        P.write_r = st < 2;
        P.want_visc_max = (st == 2 && last_visc) ? 1 : 0;
        A.s1 = as<R>(s->sets[in]);
        A.qout = as<R>(s->sets[outs[st]]);
        A.rout = st < 2 ? as<R>(s->rbuf[st]) : nullptr;
        A.ra = st == 0 ? nullptr : as<R>(s->rbuf[st - 1]);
        // every stage output is the next stage's running sum
        A.s1lo = as<R>(s->lo[in]);
        A.qoutlo = as<R>(s->lo[outs[st]]);
        P.write_qlo = 1;
    } else if (c.scheme == CNS_RK3_KUTTA) {
        // q1 = q0 + dt/2 k1; q2 = q0 + dt (2 k2 - k1); q = q0 + dt (k1 + 4 k2 + k3)/6
        const int q0 = cur, s1 = (cur + 1) % 3, s2 = (cur + 2) % 3;
        const int ins[3] = {q0, s1, s2}, outs[3] = {s1, s2, q0};
        in = ins[st];
        if (st == 0) P = stage_params(s, 0, 0, 0.5);
        else if (st == 1) P = stage_params(s, -1.0, 0, 2.0);
        else P = stage_params(s, 1.0 / 6.0, 4.0 / 6.0, 1.0 / 6.0);
        P.write_r = st < 2;
        P.want_visc_max = (st == 2 && last_visc) ? 1 : 0;
        A.qout = as<R>(s->sets[outs[st]]);
        A.s1 = as<R>(s->sets[q0]);  // the stage-0 input is q0 itself
        A.rout = st < 2 ? as<R>(s->rbuf[st]) : nullptr;
        A.ra = as<R>(s->rbuf[0]);
        A.rb = as<R>(s->rbuf[1]);
        // q1, q2 only feed the RHS; the step result carries a low part
        A.s1lo = as<R>(s->lo[q0]);
        A.qoutlo = as<R>(s->lo[outs[st]]);
        P.write_qlo = st == 2;
    } else {
        // classic RK4 with an accumulator set: acc = q0 + dt/6 k1 + dt/3 k2 + dt/3 k3,
        // q = acc + dt/6 k4.  Sets: A = current, B/D stage inputs, C accumulator.
        const int sa = cur, sb = (cur + 1) % 4, scc = (cur + 2) % 4, sd = (cur + 3) % 4;
        const int ins[4] = {sa, sb, sd, sb}, outs[4] = {sb, sd, sb, sa};
        in = ins[st];
        if (st == 0) P = stage_params(s, 0, 0, 0.5);
        else if (st == 1) P = stage_params(s, 0, 0, 0.5);
        else if (st == 2) P = stage_params(s, 0, 0, 1.0);
        else P = stage_params(s, 0, 0, 1.0 / 6.0);
        A.qout = as<R>(s->sets[outs[st]]);
        A.s1 = as<R>(s->sets[st < 3 ? sa : scc]);
        A.s1lo = as<R>(s->lo[st < 3 ? sa : scc]);
        // stage inputs only feed the RHS; the step result carries a low part
        A.qoutlo = as<R>(s->lo[outs[st]]);
        P.write_qlo = st == 3;
        if (st < 3) {
            // acc = q0 + dt/6 k1, then acc += dt/3 k2, acc += dt/3 k3
            P.write_acc = 1;
            P.v_r = st == 0 ? 1.0 / 6.0 : 1.0 / 3.0;
            A.s2 = as<R>(s->sets[st == 0 ? sa : scc]);
            A.accout = as<R>(s->sets[scc]);
            A.s2lo = as<R>(s->lo[st == 0 ? sa : scc]);
            A.accoutlo = as<R>(s->lo[scc]);
        }
        P.want_visc_max = (st == 3 && last_visc) ? 1 : 0;
    }
    A.qin = as<R>(s->sets[in]);
    A.fs = s->n;
    A.dil = as<R>(s->d_dil);
    fill_unused(A);
    return in;
}

// queue nsteps steps on the stream; outputs stay on the device
template <typename R>
static int queue_steps_t(cns_solver *s, int nsteps, bool want_par, bool want_time) {
    const cns_config &c = s->cfg;
    const long n = s->n;
    (void)n;
    for (int istep = 0; istep < nsteps; istep++) {
        StagePtrs<R> cur = ptrs_in<R>(s, s->cur);
        const bool cfl_step = (istep % c.check_cfl) == 0;
        const bool bulk_step = want_par && (istep % c.check_bulk) == 0;
        if (cfl_step) {
            const int vfs = s->visc_from_state ? 1 : 0;
            if (vfs) {
                hipLaunchKernelGGL(k_reset_visc, dim3(1), dim3(1), 0, s->stream, s->d_sc);
                LAUNCHED(s, "k_reset_visc");
            }
            hipLaunchKernelGGL(k_cfl<R>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s->stream, s->base, cur, vfs,
                               s->d_sc);
            LAUNCHED(s, "k_cfl<R>");
            hipLaunchKernelGGL(k_dt_finalize, dim3(1), dim3(1), 0, s->stream, c.cfl, s->d_sc);
            LAUNCHED(s, "k_dt_finalize");
            s->visc_from_state = false;
        }
        const int what = ((cfl_step && c.forcing_z) ? 1 : 0) | (bulk_step ? 2 : 0);
        if (what) {
            hipLaunchKernelGGL(k_partials<R>, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s->stream, s->base, cur,
                               s->d_part);
            LAUNCHED(s, "k_partials<R>");
        }
        if (what & 1) {
            hipLaunchKernelGGL(k_final, dim3(1), dim3(RED_THREADS), 0, s->stream, s->d_part, s->d_sc, 1, c.forcing_z,
                               c.len_x * c.len_y * c.len_z, (double *)nullptr, (double *)nullptr, istep);
            LAUNCHED(s, "k_final");
        }
        hipLaunchKernelGGL(k_advance_time, dim3(1), dim3(1), 0, s->stream, s->d_sc, want_time ? s->d_time : nullptr,
                           istep);
        LAUNCHED(s, "k_advance_time");
        if (what & 2) {
            hipLaunchKernelGGL(k_final, dim3(1), dim3(RED_THREADS), 0, s->stream, s->d_part, s->d_sc, 2, c.forcing_z,
                               c.len_x * c.len_y * c.len_z, s->d_par1, s->d_par2, istep);
            LAUNCHED(s, "k_final");
        }
        const bool need_visc = ((istep + 1) % c.check_cfl == 0) || (istep == nsteps - 1);
        for (int st = 0; st < n_stages(s); st++) {
            Params P;
            StagePtrs<R> A;
            const int in = stage_setup<R>(s, s->cur, st, need_visc, P, A);
            int rc;
            if ((rc = run_stage<R>(s, in, P, A))) return rc;
        }
        s->cur = next_cur(s, s->cur);
    }
    return CNS_OK;
}

static int queue_steps(cns_solver *s, int nsteps, bool want_par, bool want_time) {
    return s->fp32 ? queue_steps_t<float>(s, nsteps, want_par, want_time)
                   : queue_steps_t<double>(s, nsteps, want_par, want_time);
}

static int ensure_par(cns_solver *s, int nsteps) {
    if (nsteps <= s->par_cap) return CNS_OK;
    // the arrays may still be in use by queued steps
    HIPCHK(s, hipStreamSynchronize(s->stream));
    dev_free(s->dev, s->d_par1);
    dev_free(s->dev, s->d_par2);
    dev_free(s->dev, s->d_time);
    s->d_par1 = s->d_par2 = s->d_time = nullptr;
    HIPCHK(s, dev_alloc(s->dev, (void **)&s->d_par1, nsteps * sizeof(double)));
    HIPCHK(s, dev_alloc(s->dev, (void **)&s->d_par2, nsteps * sizeof(double)));
    HIPCHK(s, dev_alloc(s->dev, (void **)&s->d_time, nsteps * sizeof(double)));
    s->par_cap = nsteps;
    return CNS_OK;
}

extern "C" int cns_step(cns_solver *s, int nsteps, double *kin, double *enst, double *time) {
    if (!s) return fail(nullptr, CNS_ERR_ARG, "null solver");
    if (nsteps < 0) return fail(s, CNS_ERR_ARG, "cns_step: nsteps < 0");
    if (!s->have_state) return fail(s, CNS_ERR_STATE, "cns_step: no state set");
    if (s->in_batch) return fail(s, CNS_ERR_STATE, "cns_step: the solver is advanced by a batch");
    s->bsteps = 0;
    if (nsteps == 0) return CNS_OK;
    HIPCHK(s, hipSetDevice(s->dev));
    const bool want_par = kin || enst;
    int rc;
    if ((rc = ensure_par(s, nsteps))) return rc;
    if (want_par) {
        const double nan = std::nan("");
        hipLaunchKernelGGL(k_fill, dim3(64), dim3(256), 0, s->stream, s->d_par1, (long)nsteps, nan);
        hipLaunchKernelGGL(k_fill, dim3(64), dim3(256), 0, s->stream, s->d_par2, (long)nsteps, nan);
        LAUNCHED(s, "k_fill");
    }
    if ((rc = queue_steps(s, nsteps, want_par, time != nullptr))) return rc;
    // kin, enst and time may be host arrays or device arrays (the reference
    // passes device arrays, src/main.h:35): the runtime takes the direction
    // from the addresses
    const size_t b = nsteps * sizeof(double);
    if (kin) HIPCHK(s, hipMemcpyAsync(kin, s->d_par1, b, hipMemcpyDefault, s->stream));
    if (enst) HIPCHK(s, hipMemcpyAsync(enst, s->d_par2, b, hipMemcpyDefault, s->stream));
    if (time) HIPCHK(s, hipMemcpyAsync(time, s->d_time, b, hipMemcpyDefault, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return CNS_OK;
}

extern "C" int cns_set_pencil(cns_solver *s, int pencil) {
    if (!s) return fail(nullptr, CNS_ERR_ARG, "null solver");
    if (s->in_batch) return fail(s, CNS_ERR_STATE, "cns_set_pencil: the solver is advanced by a batch");
    if (pencil < 1) return fail(s, CNS_ERR_ARG, "cns_set_pencil: pencil must be >= 1");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    set_pencil(s, pencil);
    return CNS_OK;
}

extern "C" int cns_tune(cns_solver *s, int *pencil) {
    if (!s) return fail(nullptr, CNS_ERR_ARG, "null solver");
    if (!s->have_state) return fail(s, CNS_ERR_STATE, "cns_tune: no state set");
    if (s->in_batch) return fail(s, CNS_ERR_STATE, "cns_tune: the solver is advanced by a batch");
    HIPCHK(s, hipSetDevice(s->dev));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    // snapshot of everything a step changes: the state set (and its low
    // part), the device scalars and the host step control
    const size_t bytes = 5 * (size_t)s->n * s->esz;
    const int cur0 = s->cur;
    const long bsteps0 = s->bsteps;
    const bool vfs0 = s->visc_from_state;
    std::vector<char> hq(bytes), hl(s->comp ? bytes : 0);
    Scalars sc0{};
    HIPCHK(s, hipMemcpy(hq.data(), s->sets[cur0], bytes, hipMemcpyDeviceToHost));
    if (s->comp) HIPCHK(s, hipMemcpy(hl.data(), s->lo[cur0], bytes, hipMemcpyDeviceToHost));
    HIPCHK(s, hipMemcpy(&sc0, s->d_sc, sizeof(sc0), hipMemcpyDeviceToHost));
    std::vector<int> cand;
    for (int L : {8, 16, 32, 64, 128, 256})
        if (L <= s->cfg.nz) cand.push_back(L);
    if (std::find(cand.begin(), cand.end(), s->zc) == cand.end()) cand.push_back(s->zc);
    const int steps = (int)std::max(1L, std::min(4L, 200000000L / s->n));
    hipEvent_t a, b;
    HIPCHK(s, hipEventCreate(&a));
    HIPCHK(s, hipEventCreate(&b));
    int best = s->zc, rc = CNS_OK;
    float bt = 1e30f;
    for (int L : cand) {
        set_pencil(s, L);
        if ((rc = queue_steps(s, 1, false, false))) break;
        if (hipEventRecord(a, s->stream) != hipSuccess) break;
        if ((rc = queue_steps(s, steps, false, false))) break;
        float t = 0;
        if (hipEventRecord(b, s->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess ||
            hipEventElapsedTime(&t, a, b) != hipSuccess) {
            rc = fail(s, CNS_ERR_HIP, "cns_tune: event timing failed");
            break;
        }
        if (t < bt) {
            bt = t;
            best = L;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    // restore the snapshot, bit for bit
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(s->sets[cur0], hq.data(), bytes, hipMemcpyHostToDevice));
    if (s->comp) HIPCHK(s, hipMemcpy(s->lo[cur0], hl.data(), bytes, hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->d_sc, &sc0, sizeof(sc0), hipMemcpyHostToDevice));
    s->cur = cur0;
    s->bsteps = bsteps0;
    s->visc_from_state = vfs0;
    set_pencil(s, best);
    if (rc) return rc;
    if (pencil) *pencil = best;
    return CNS_OK;
}

// Events of one timed region, destroyed on every return path.
namespace {
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};
}  // namespace

// Times nsteps steps exactly as cns_step queues them with all outputs: CFL
// and forcing checks, the bulk statistics every check_bulk steps
// (src/cuda_main.cu:109-186) and the time array; only the copies of those
// small arrays to the host are left out.
extern "C" int cns_time_steps(cns_solver *s, int nsteps, double *ms) {
    if (!s || !ms) return fail(s, CNS_ERR_ARG, "cns_time_steps: null argument");
    if (nsteps < 1) return fail(s, CNS_ERR_ARG, "cns_time_steps: nsteps < 1");
    if (!s->have_state) return fail(s, CNS_ERR_STATE, "cns_time_steps: no state set");
    if (s->in_batch) return fail(s, CNS_ERR_STATE, "cns_time_steps: the solver is advanced by a batch");
    s->bsteps = 0;
    HIPCHK(s, hipSetDevice(s->dev));
    int rc;
    if ((rc = ensure_par(s, nsteps))) return rc;
    EventPair ev;
    HIPCHK(s, hipEventCreate(&ev.a));
    HIPCHK(s, hipEventCreate(&ev.b));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipEventRecord(ev.a, s->stream));
    const double nan = std::nan("");
    hipLaunchKernelGGL(k_fill, dim3(64), dim3(256), 0, s->stream, s->d_par1, (long)nsteps, nan);
    hipLaunchKernelGGL(k_fill, dim3(64), dim3(256), 0, s->stream, s->d_par2, (long)nsteps, nan);
    LAUNCHED(s, "k_fill");
    if ((rc = queue_steps(s, nsteps, true, true))) return rc;
    HIPCHK(s, hipEventRecord(ev.b, s->stream));
    HIPCHK(s, hipEventSynchronize(ev.b));
    float f = 0;
    HIPCHK(s, hipEventElapsedTime(&f, ev.a, ev.b));
    *ms = f;
    return CNS_OK;
}

// Natural cubic spline through (xk, yk) (zero second derivative at both
// ends), the interpolation of calculateSponge (src/sponge.cu:258-313): the
// knots' second derivatives from the tridiagonal system, then the cubic of
// the interval holding each abscissa (the end intervals' cubics continue
// outside the knots).
namespace {
struct NaturalSpline {
    const double *xk, *yk;
    int n;
    std::vector<double> m;
    NaturalSpline(const double *x, const double *y, int count) : xk(x), yk(y), n(count), m(count, 0.0) {
        if (n < 3) return;
        const int k = n - 2;
        std::vector<double> cp(k), dp(k);
        for (int i = 0; i < k; i++) {
            const double h0 = xk[i + 1] - xk[i], h1 = xk[i + 2] - xk[i + 1];
            const double b = 2.0 * (h0 + h1);
            const double d = 6.0 * ((yk[i + 2] - yk[i + 1]) / h1 - (yk[i + 1] - yk[i]) / h0);
            const double den = i ? b - h0 * cp[i - 1] : b;
            cp[i] = i < k - 1 ? h1 / den : 0.0;
            dp[i] = i ? (d - h0 * dp[i - 1]) / den : d / den;
        }
        m[k] = dp[k - 1];
        for (int i = k - 2; i >= 0; i--) m[i + 1] = dp[i] - cp[i] * m[i + 2];
    }
    double operator()(double x) const {
        int lo = 0, hi = n - 1;  // largest knot <= x, clamped to the first and last interval
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (xk[mid] <= x)
                lo = mid;
            else
                hi = mid;
        }
        const double h = xk[lo + 1] - xk[lo];
        const double a = (xk[lo + 1] - x) / h, b = (x - xk[lo]) / h;
        return a * yk[lo] + b * yk[lo + 1] + ((a * a * a - a) * m[lo] + (b * b * b - b) * m[lo + 1]) * h * h / 6.0;
    }
};
}  // namespace

extern "C" int cns_sponge_from_profile(int n, const double *yp, const double *rp, const double *up, const double *wp,
                                       int mx, const double *x, int mz, const double *z, double len_x, double len_z,
                                       double reynolds, double gamma, double mach, const double *zones,
                                       double *sigma_x, double *sigma_z, double *rref, double *ruref, double *rvref,
                                       double *rwref, double *reref) {
    if (n < 2 || mx < 1 || mz < 1 || !yp || !rp || !up || !wp || !x || !z || !sigma_x || !sigma_z || !rref ||
        !ruref || !rvref || !rwref || !reref)
        return fail(nullptr, CNS_ERR_ARG, "cns_sponge_from_profile: bad argument");
    for (int i = 1; i < n; i++)
        if (!(yp[i] > yp[i - 1])) return fail(nullptr, CNS_ERR_ARG, "cns_sponge_from_profile: profile abscissae must increase");
    // src/sponge.h:4-17: (strength, length, exponent) of the top, inlet and outlet zones
    static const double ref_zones[9] = {1.0, 1.0, 2.0, 0.5, 20.0, 2.0, 0.5, 20.0, 2.0};
    const double *zn = zones ? zones : ref_zones;
    // src/sponge.cu:112-129; the outlet zone wins where inlet and outlet overlap
    for (int i = 0; i < mx; i++) {
        sigma_x[i] = 0.0;
        if (zn[1] > 0.0 && x[i] >= len_x - zn[1]) sigma_x[i] = zn[0] * std::pow((x[i] - (len_x - zn[1])) / zn[1], zn[2]);
    }
    for (int k = 0; k < mz; k++) {
        sigma_z[k] = 0.0;
        if (zn[4] > 0.0 && z[k] <= zn[4]) sigma_z[k] = zn[3] * std::pow((zn[4] - z[k]) / zn[4], zn[5]);
        if (zn[7] > 0.0 && z[k] >= len_z - zn[7]) sigma_z[k] = zn[6] * std::pow((z[k] - (len_z - zn[7])) / zn[7], zn[8]);
    }
    // src/sponge.cu:158-168: profile read at x / sqrt(1 + z/Re), wall-normal
    // velocity over (scale Re), pressure Rgas; conserved variables
    const NaturalSpline sr(yp, rp, n), su(yp, up, n), sw(yp, wp, n);
    const double rgas = 1.0 / (gamma * mach * mach);
    for (int k = 0; k < mz; k++) {
        const double scale = std::sqrt(1.0 + z[k] / reynolds);
        for (int i = 0; i < mx; i++) {
            const size_t g = i + (size_t)k * mx;
            const double xs = x[i] / scale;
            const double r = sr(xs), u = su(xs) / (scale * reynolds), w = sw(xs);
            rref[g] = r;
            ruref[g] = r * u;
            rvref[g] = 0.0;
            rwref[g] = r * w;
            reref[g] = rgas / (gamma - 1.0) + r * 0.5 * (u * u + w * w);
        }
    }
    return CNS_OK;
}

extern "C" int cns_set_sponge(cns_solver *s, const double *sigma_x, const double *sigma_z, const double *rref,
                              const double *ruref, const double *rvref, const double *rwref, const double *reref) {
    if (!s || !sigma_x || !sigma_z || !rref || !ruref || !rvref || !rwref || !reref)
        return fail(s, CNS_ERR_ARG, "cns_set_sponge: null argument");
    HIPCHK(s, hipSetDevice(s->dev));
    const int mx = s->cfg.nx, mz = s->cfg.nz;
    HIPCHK(s, hipMemcpy(s->d_spx, sigma_x, mx * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(s, hipMemcpy(s->d_spz, sigma_z, mz * sizeof(double), hipMemcpyHostToDevice));
    const double *src[5] = {rref, ruref, rvref, rwref, reref};
    for (int q = 0; q < 5; q++)
        HIPCHK(s, hipMemcpy(s->d_ref + (size_t)q * mx * mz, src[q], (size_t)mx * mz * sizeof(double),
                            hipMemcpyHostToDevice));
    return CNS_OK;
}

extern "C" int cns_get_scalars(cns_solver *s, double *dt, double *dpdz, double *time) {
    if (!s) return fail(nullptr, CNS_ERR_ARG, "null solver");
    HIPCHK(s, hipSetDevice(s->dev));
    Scalars h{};
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(&h, s->d_sc, sizeof(h), hipMemcpyDeviceToHost));
    if (dt) *dt = h.dt;
    if (dpdz) *dpdz = h.dpdz;
    if (time) *time = h.time;
    return CNS_OK;
}

extern "C" int cns_sync(cns_solver *s) {
    if (!s) return fail(nullptr, CNS_ERR_ARG, "null solver");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return CNS_OK;
}

extern "C" int cns_profile_enable(cns_solver *s, int enable) {
    if (!s) return fail(nullptr, CNS_ERR_ARG, "null solver");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    s->prof = enable != 0;
    s->prof_n = 0;
    return CNS_OK;
}

extern "C" int cns_profile_read(cns_solver *s, double *total_ms, int *launches) {
    if (!s || !total_ms || !launches) return fail(s, CNS_ERR_ARG, "cns_profile_read: null argument");
    HIPCHK(s, hipStreamSynchronize(s->stream));
    double tot = 0.0;
    for (int k = 0; k < s->prof_n; k++) {
        float f = 0;
        HIPCHK(s, hipEventElapsedTime(&f, s->ev[2 * k], s->ev[2 * k + 1]));
        tot += f;
    }
    *total_ms = tot;
    *launches = s->prof_n;
    return CNS_OK;
}

extern "C" int cns_mem_info(int device, uint64_t *free_bytes, uint64_t *total_bytes) {
    if (!free_bytes || !total_bytes) return fail(nullptr, CNS_ERR_ARG, "cns_mem_info: null argument");
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) return fail(nullptr, CNS_ERR_HIP, "cns_mem_info: no HIP device");
    size_t f = 0, t = 0;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipMemGetInfo(&f, &t);
    (void)hipSetDevice(prev);
    if (e != hipSuccess) return fail(nullptr, CNS_ERR_HIP, std::string("cns_mem_info: ") + hipGetErrorString(e));
    *free_bytes = f;
    *total_bytes = t;
    return CNS_OK;
}

extern "C" int cns_device_bytes(const cns_solver *s, uint64_t *bytes) {
    if (!s || !bytes) return fail(nullptr, CNS_ERR_ARG, "cns_device_bytes: null argument");
    *bytes = s->dev_bytes;
    return CNS_OK;
}

// ---------------------------------------------------------------------------
// Diagnostics of the current state (cns_diag.hpp)
// ---------------------------------------------------------------------------

// stored state set and its low part (null unless compensated FP32)
template <typename R>
static void state_arrays(const cns_solver *s, const R *&q, const R *&lo) {
    q = as<R>(s->sets[s->cur]);
    lo = (s->fp32 && s->comp) ? as<R>(s->lo[s->cur]) : nullptr;
}

static int diag_ready(cns_solver *s, const char *what) {
    if (s->in_batch) return fail(s, CNS_ERR_STATE, std::string(what) + ": the solver is advanced by a batch");
    if (!s->have_state) return fail(s, CNS_ERR_STATE, std::string(what) + ": no state set");
    HIPCHK(s, hipSetDevice(s->dev));
    return CNS_OK;
}

template <typename R>
static int xprof_t(cns_solver *s, double *prof) {
    const int mx = s->cfg.nx, my = s->cfg.ny, mz = s->cfg.nz;
    const long rows = (long)my * mz;
    if (!s->d_prof) {
        int rc = balloc(s, (void **)&s->d_prof, 10 * (size_t)mx * sizeof(double));
        if (rc) return rc;
    }
    const R *q, *lo;
    state_arrays<R>(s, q, lo);
    // chunk partials in the first rhs set (free between steps):
    // nch * 5 * mx doubles within its 5 * n * esz bytes
    double *part = static_cast<double *>(s->rbuf[0]);
    const int nch = (int)std::max(1L, std::min(256L, rows * (long)s->esz / 8));
    const dim3 grid((mx + 255) / 256, nch), fin((mx + 255) / 256);
    for (int pass = 0; pass < 2; pass++) {
        hipLaunchKernelGGL(k_xprof_sums<R>, grid, dim3(256), 0, s->stream, mx, rows, (long)s->n, q, lo, s->rho0, s->E0,
                           pass, (const double *)s->d_prof, part, nch);
        LAUNCHED(s, "k_xprof_sums");
        hipLaunchKernelGGL(k_xprof_final, fin, dim3(256), 0, s->stream, mx, my, mz, (const double *)part, nch, pass,
                           s->d_prof);
        LAUNCHED(s, "k_xprof_final");
    }
    std::copy(s->x.begin(), s->x.end(), prof);
    HIPCHK(s, hipMemcpyAsync(prof + mx, s->d_prof, 10 * (size_t)mx * sizeof(double), hipMemcpyDeviceToHost,
                             s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return CNS_OK;
}

extern "C" int cns_channel_profiles(cns_solver *s, double *prof) {
    if (!s || !prof) return fail(s, CNS_ERR_ARG, "cns_channel_profiles: null argument");
    int rc = diag_ready(s, "cns_channel_profiles");
    if (rc) return rc;
    return s->fp32 ? xprof_t<float>(s, prof) : xprof_t<double>(s, prof);
}

template <typename R, int S>
static void launch_ret(cns_solver *s, const R *q, const R *lo) {
    const double muw = std::pow(1.0, s->cfg.visc_exp) / s->cfg.reynolds;
    hipLaunchKernelGGL((k_ret<R, S>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s->stream, s->cfg.nx,
                       (long)s->cfg.ny * s->cfg.nz, (long)s->n, q, lo, s->rho0, s->E0, s->dx, s->xp[0], muw, s->cF[0],
                       s->cF[1], s->cF[2], s->cF[3], s->d_part);
}

template <typename R>
static int ret_t(cns_solver *s, double *re_tau) {
    const R *q, *lo;
    state_arrays<R>(s, q, lo);
    switch (s->S) {
        case 1: launch_ret<R, 1>(s, q, lo); break;
        case 2: launch_ret<R, 2>(s, q, lo); break;
        case 3: launch_ret<R, 3>(s, q, lo); break;
        default: launch_ret<R, 4>(s, q, lo); break;
    }
    LAUNCHED(s, "k_ret");
    std::vector<double> part(RED_BLOCKS);
    HIPCHK(s, hipMemcpyAsync(part.data(), s->d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    double a = 0.0;
    for (double v : part) a += v;
    *re_tau = a / s->cfg.ny / s->cfg.nz;
    return CNS_OK;
}

extern "C" int cns_friction_reynolds(cns_solver *s, double *re_tau) {
    if (!s || !re_tau) return fail(s, CNS_ERR_ARG, "cns_friction_reynolds: null argument");
    int rc = diag_ready(s, "cns_friction_reynolds");
    if (rc) return rc;
    return s->fp32 ? ret_t<float>(s, re_tau) : ret_t<double>(s, re_tau);
}

template <typename R, int SV>
static void launch_stats(cns_solver *s, const R *q, const R *lo) {
    hipLaunchKernelGGL((k_flow_stats<R, SV>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s->stream, s->base,
                       (long)s->n, q, lo, s->cfg.periodic_x ? 0 : 1, s->cfg.nonuniform_x ? 1 : 0, s->d_part);
}

template <typename R>
static int stats_t(cns_solver *s, double *out) {
    const R *q, *lo;
    state_arrays<R>(s, q, lo);
    switch (s->SV) {
        case 1: launch_stats<R, 1>(s, q, lo); break;
        case 2: launch_stats<R, 2>(s, q, lo); break;
        case 3: launch_stats<R, 3>(s, q, lo); break;
        default: launch_stats<R, 4>(s, q, lo); break;
    }
    LAUNCHED(s, "k_flow_stats");
    std::vector<double> part(4 * RED_BLOCKS);
    HIPCHK(s, hipMemcpyAsync(part.data(), s->d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    double sum[4] = {0, 0, 0, 0};
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < RED_BLOCKS; b++) sum[a] += part[a * RED_BLOCKS + b];
    const double vol = s->cfg.len_x * s->cfg.len_y * s->cfg.len_z;
    out[0] = sum[1] / sum[0];
    out[1] = sum[2] / sum[0];
    out[2] = sum[3] / vol;
    out[3] = sum[0] / vol;
    return CNS_OK;
}

// flow statistics of the current state into dst[0..3] on the device
template <typename R>
static int queue_stats(cns_solver *s, double *dst) {
    const R *q, *lo;
    state_arrays<R>(s, q, lo);
    switch (s->SV) {
        case 1: launch_stats<R, 1>(s, q, lo); break;
        case 2: launch_stats<R, 2>(s, q, lo); break;
        case 3: launch_stats<R, 3>(s, q, lo); break;
        default: launch_stats<R, 4>(s, q, lo); break;
    }
    LAUNCHED(s, "k_flow_stats");
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(64), 0, s->stream, (const double *)s->d_part,
                       s->cfg.len_x * s->cfg.len_y * s->cfg.len_z, dst);
    LAUNCHED(s, "k_stats_final");
    return CNS_OK;
}

extern "C" int cns_step_stats(cns_solver *s, int nsteps, int every, double *stats) {
    if (!s || !stats) return fail(s, CNS_ERR_ARG, "cns_step_stats: null argument");
    if (nsteps < 0 || every < 1 || nsteps % every != 0)
        return fail(s, CNS_ERR_ARG, "cns_step_stats: n_steps must be a non-negative multiple of every >= 1");
    int rc = diag_ready(s, "cns_step_stats");
    if (rc) return rc;
    const int nsamp = nsteps / every;
    if (nsamp == 0) return CNS_OK;
    if (4 * nsamp > s->stats_cap) {
        if (s->d_stats) {
            HIPCHK(s, hipStreamSynchronize(s->stream));
            dev_free(s->dev, s->d_stats);
            s->dev_bytes -= (uint64_t)s->stats_cap * sizeof(double);
            s->d_stats = nullptr;
            s->stats_cap = 0;
        }
        if ((rc = balloc(s, (void **)&s->d_stats, 4 * (size_t)nsamp * sizeof(double)))) return rc;
        s->stats_cap = 4 * nsamp;
    }
    s->bsteps = 0;
    for (int k = 0; k < nsamp; k++) {
        if ((rc = queue_steps(s, every, false, false))) return rc;
        if ((rc = s->fp32 ? queue_stats<float>(s, s->d_stats + 4 * k) : queue_stats<double>(s, s->d_stats + 4 * k)))
            return rc;
    }
    HIPCHK(s, hipMemcpyAsync(stats, s->d_stats, 4 * (size_t)nsamp * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(s, hipStreamSynchronize(s->stream));
    return CNS_OK;
}

extern "C" int cns_flow_stats(cns_solver *s, double *out) {
    if (!s || !out) return fail(s, CNS_ERR_ARG, "cns_flow_stats: null argument");
    int rc = diag_ready(s, "cns_flow_stats");
    if (rc) return rc;
    return s->fp32 ? stats_t<float>(s, out) : stats_t<double>(s, out);
}

// ---------------------------------------------------------------------------
// Field files (src/comm.cpp:205-279, src/init.cpp:13-30,78-88) and control
// ---------------------------------------------------------------------------

static const char FIELD_NAMES[5] = {'r', 'u', 'v', 'w', 'e'};

static std::string field_path(const char *dir, char name, int timestep) {
    char f[32];
    std::snprintf(f, sizeof(f), "%c.%07d.bin", name, timestep);
    return std::string(dir) + "/" + f;
}

// writes count doubles to path; false on any failure (the message in err)
static bool write_f64(const std::string &path, const double *a, size_t count, std::string &err) {
    std::FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        err = "cannot open " + path + " for writing: " + std::strerror(errno);
        return false;
    }
    const size_t got = std::fwrite(a, sizeof(double), count, f);
    const bool closed = std::fclose(f) == 0;
    if (got != count || !closed) {
        err = "short write to " + path;
        return false;
    }
    return true;
}

// reads exactly count doubles from path, which must hold no more and no less
static bool read_f64(const std::string &path, double *a, size_t count, std::string &err) {
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open " + path + ": " + std::strerror(errno);
        return false;
    }
    bool ok = std::fseek(f, 0, SEEK_END) == 0;
    const long size = ok ? std::ftell(f) : -1;
    ok = ok && size >= 0 && (size_t)size == count * sizeof(double) && std::fseek(f, 0, SEEK_SET) == 0;
    if (!ok) {
        std::fclose(f);
        err = path + " holds " + std::to_string(size) + " bytes, expected " +
              std::to_string(count * sizeof(double)) + " (float64 of the whole grid)";
        return false;
    }
    const size_t got = std::fread(a, sizeof(double), count, f);
    std::fclose(f);
    if (got != count) {
        err = "short read from " + path;
        return false;
    }
    return true;
}

extern "C" int cns_write_fields(cns_solver *s, const char *dir, int timestep) {
    if (!s || !dir) return fail(s, CNS_ERR_ARG, "cns_write_fields: null argument");
    if (timestep < 0 || timestep > 9999999) return fail(s, CNS_ERR_ARG, "cns_write_fields: timestep outside 0..9999999");
    const size_t n = (size_t)s->n;
    std::vector<double> q(5 * n);
    int rc = cns_get_state(s, q.data(), q.data() + n, q.data() + 2 * n, q.data() + 3 * n, q.data() + 4 * n);
    if (rc) return rc;
    std::string err;
    for (int k = 0; k < 5; k++)
        if (!write_f64(field_path(dir, FIELD_NAMES[k], timestep), q.data() + k * n, n, err))
            return fail(s, CNS_ERR_IO, "cns_write_fields: " + err);
    return CNS_OK;
}

extern "C" int cns_read_fields(cns_solver *s, const char *dir, int timestep) {
    if (!s || !dir) return fail(s, CNS_ERR_ARG, "cns_read_fields: null argument");
    if (timestep < 0 || timestep > 9999999) return fail(s, CNS_ERR_ARG, "cns_read_fields: timestep outside 0..9999999");
    if (s->in_batch) return fail(s, CNS_ERR_STATE, "cns_read_fields: the solver is advanced by a batch");
    const size_t n = (size_t)s->n;
    std::vector<double> q(5 * n);
    std::string err;
    for (int k = 0; k < 5; k++)
        if (!read_f64(field_path(dir, FIELD_NAMES[k], timestep), q.data() + k * n, n, err))
            return fail(s, CNS_ERR_IO, "cns_read_fields: " + err);
    return cns_set_state(s, q.data(), q.data() + n, q.data() + 2 * n, q.data() + 3 * n, q.data() + 4 * n);
}

extern "C" int cns_write_grid(const cns_solver *s, const char *dir) {
    if (!s || !dir) return fail(nullptr, CNS_ERR_ARG, "cns_write_grid: null argument");
    std::string err;
    const std::string d(dir);
    if (!write_f64(d + "/x.bin", s->x.data(), s->x.size(), err) ||
        !write_f64(d + "/y.bin", s->y.data(), s->y.size(), err) ||
        !write_f64(d + "/z.bin", s->z.data(), s->z.size(), err))
        return fail(const_cast<cns_solver *>(s), CNS_ERR_IO, "cns_write_grid: " + err);
    return CNS_OK;
}

extern "C" int cns_get_control(cns_solver *s, cns_control *c) {
    if (!s || !c) return fail(s, CNS_ERR_ARG, "cns_get_control: null argument");
    if (s->in_batch) return fail(s, CNS_ERR_STATE, "cns_get_control: the solver is advanced by a batch");
    HIPCHK(s, hipSetDevice(s->dev));
    Scalars h{};
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(&h, s->d_sc, sizeof(h), hipMemcpyDeviceToHost));
    std::memset(c, 0, sizeof(*c));
    c->time = h.time;
    c->dt = h.dt;
    c->dpdz = h.dpdz;
    // the maximum is kept as the bits of a non-negative double
    std::memcpy(&c->visc_max, &h.visc_max, sizeof(double));
    c->steps = s->bsteps;
    c->visc_from_state = s->visc_from_state ? 1 : 0;
    return CNS_OK;
}

extern "C" int cns_set_control(cns_solver *s, const cns_control *c) {
    if (!s || !c) return fail(s, CNS_ERR_ARG, "cns_set_control: null argument");
    if (s->in_batch) return fail(s, CNS_ERR_STATE, "cns_set_control: the solver is advanced by a batch");
    if (!s->have_state) return fail(s, CNS_ERR_STATE, "cns_set_control: no state set");
    if (!(c->dt >= 0.0) || !(c->visc_max >= 0.0) || c->steps < 0 || !std::isfinite(c->time))
        return fail(s, CNS_ERR_ARG, "cns_set_control: dt, visc_max and steps must be non-negative, time finite");
    HIPCHK(s, hipSetDevice(s->dev));
    Scalars h{};
    HIPCHK(s, hipStreamSynchronize(s->stream));
    HIPCHK(s, hipMemcpy(&h, s->d_sc, sizeof(h), hipMemcpyDeviceToHost));
    h.time = c->time;
    h.dt = c->dt;
    h.dpdz = c->dpdz;
    std::memcpy(&h.visc_max, &c->visc_max, sizeof(double));
    HIPCHK(s, hipMemcpy(s->d_sc, &h, sizeof(h), hipMemcpyHostToDevice));
    s->bsteps = c->steps;
    s->visc_from_state = c->visc_from_state != 0;
    return CNS_OK;
}

// ---------------------------------------------------------------------------
// Batched advance (include/cns.h: cns_batch_*)
// ---------------------------------------------------------------------------

namespace {

struct BatchGroup {
    std::vector<cns_solver *> m;
    bool fp32 = false, wallx = false, nonuni = false, unitv = false, forcing = false;
    int S = 0, SV = 0, scheme = 0, nst = 0;
    int pencil = 64, grid = 0, remap = 1, nwork = 0;
    void *d_jobs = nullptr;  // nst arrays of m.size() jobs
    size_t job_size = 0;
};

}  // namespace

struct cns_batch {
    std::string err;
    int dev = 0, cus = 256;
    hipStream_t stream = nullptr;
    std::vector<cns_solver *> members;
    std::vector<int> cur0;     // each member's state set when the batch was made
    std::vector<long> phase0;  // each member's batch step count when the batch was made
    std::vector<BatchGroup> groups;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    long steps = 0;
    int pending = 0;  // steps launched by cns_batch_launch and not yet waited for
    bool sync_check = false;
};

#define BCHK(b, expr)                                                                                     \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) {                                                                           \
            (b)->err = std::string(#expr) + ": " + hipGetErrorString(e_);                                 \
            return CNS_ERR_HIP;                                                                           \
        }                                                                                                 \
    } while (0)

static int bfail(cns_batch *b, int code, const std::string &msg) {
    if (b)
        b->err = msg;
    else
        g_err = msg;
    return code;
}

// z chunks of member m under pencil length L (at least the window depth)
static void member_geometry(const cns_solver *m, int L, int &zc, int &nzc, int &nwork) {
    const int mz = m->cfg.nz;
    const int lo = std::min(2 * m->S, mz);
    zc = std::max(std::min(L, mz), lo);
    nzc = (mz + zc - 1) / zc;
    nwork = m->ntx * m->nty * nzc;
}

// Jobs of every stage of group g under pencil L.  scratch: the tuning form,
// whose outputs all go to the member's first rhs buffer and which leaves the
// CFL accumulator alone.
template <typename R>
static int upload_jobs(cns_batch *b, BatchGroup &g, int L, bool scratch) {
    const size_t nm = g.m.size();
    std::vector<Job<R>> h(nm * g.nst);
    int first_total = 0;
    for (int st = 0; st < g.nst; st++) {
        int first = 0;
        for (size_t k = 0; k < nm; k++) {
            cns_solver *m = g.m[k];
            const size_t mi = std::find(b->members.begin(), b->members.end(), m) - b->members.begin();
            const int c0 = b->cur0[mi];
            // A[p] serves steps whose count has parity p; the member's state is
            // in set c0 at its current count
            const int p0 = (int)(b->phase0[mi] & 1);
            Job<R> &J = h[st * nm + k];
            std::memset(&J, 0, sizeof(J));
            Params P1;
            stage_setup<R>(m, c0, st, true, J.P, J.A[p0]);
            stage_setup<R>(m, next_cur(m, c0), st, true, P1, J.A[1 - p0]);
            int zc, nzc, nw;
            member_geometry(m, L, zc, nzc, nw);
            J.P.zc = zc;
            J.P.nzc = nzc;
            J.P.xcd_remap = 0;
            if (scratch) {
                J.P.want_visc_max = 0;
                for (int p = 0; p < 2; p++) {
                    R *sb = as<R>(m->rbuf[0]);
                    J.A[p].qout = J.A[p].rout = J.A[p].accout = J.A[p].qoutlo = J.A[p].accoutlo = sb;
                }
            }
            J.dil = as<R>(m->d_dil);
            J.sc = m->d_sc;
            J.ctl = m->d_ctl;
            J.part = m->d_part;
            J.first = first;
            J.nwork = nw;
            first += nw;
        }
        first_total = first;
    }
    g.nwork = first_total;
    g.job_size = sizeof(Job<R>);
    if (!g.d_jobs) BCHK(b, dev_alloc(b->dev, &g.d_jobs, h.size() * sizeof(Job<R>)));
    BCHK(b, hipMemcpyAsync(g.d_jobs, h.data(), h.size() * sizeof(Job<R>), hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    return CNS_OK;
}

static int upload(cns_batch *b, BatchGroup &g, int L, bool scratch) {
    return g.fp32 ? upload_jobs<float>(b, g, L, scratch) : upload_jobs<double>(b, g, L, scratch);
}

static int blaunched(cns_batch *b, const char *what) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && b->sync_check) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        b->err = std::string(what) + ": " + hipGetErrorString(e);
        return CNS_ERR_HIP;
    }
    return CNS_OK;
}
#define BLAUNCHED(b, what)                  \
    do {                                    \
        int rc_ = blaunched((b), (what));   \
        if (rc_) return rc_;                \
    } while (0)

// dilatation + RHS launches of stage st of group g
template <typename R>
static int queue_stage(cns_batch *b, const BatchGroup &g, int st) {
    const int nm = (int)g.m.size();
    const Job<R> *jobs = static_cast<const Job<R> *>(g.d_jobs) + (size_t)st * nm;
    // grid 0: one workgroup per work item (the dispatcher balances)
    const int grid = g.grid > 0 ? std::max(1, std::min(g.grid, g.nwork)) : std::max(1, g.nwork);
    BCHK(b, launch_dil_batch<R>(g.SV, g.wallx, g.nonuni, g.unitv, grid, b->stream, jobs, nm, g.nwork, g.remap));
    BLAUNCHED(b, "k_dil_batch");
    // three update operands only in the last Kutta stage
    const bool k3 = g.scheme == CNS_RK3_KUTTA && st == 2;
    BCHK(b, launch_rhs_batch<R>(g.S, g.SV, g.wallx, g.nonuni, g.unitv, k3, grid, b->stream, jobs, nm, g.nwork,
                                g.remap));
    BLAUNCHED(b, "k_rhs_batch");
    return CNS_OK;
}

// one step of group g: control, then the stages
template <typename R>
static int queue_group_step(cns_batch *b, const BatchGroup &g) {
    const int nm = (int)g.m.size();
    const Job<R> *jobs = static_cast<const Job<R> *>(g.d_jobs);
    const int lanes = (nm + 63) / 64;
    hipLaunchKernelGGL(k_step_begin<R>, dim3(lanes), dim3(64), 0, b->stream, jobs, nm);
    BLAUNCHED(b, "k_step_begin");
    hipLaunchKernelGGL(k_cfl_batch<R>, dim3(256, nm), dim3(RED_THREADS), 0, b->stream, jobs);
    BLAUNCHED(b, "k_cfl_batch");
    if (g.forcing) {
        hipLaunchKernelGGL(k_partials_batch<R>, dim3(RED_BLOCKS, nm), dim3(RED_THREADS), 0, b->stream, jobs);
        BLAUNCHED(b, "k_partials_batch");
    }
    hipLaunchKernelGGL(k_step_mid<R>, dim3(nm), dim3(RED_THREADS), 0, b->stream, jobs);
    BLAUNCHED(b, "k_step_mid");
    for (int st = 0; st < g.nst; st++) {
        int rc = queue_stage<R>(b, g, st);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_step_end<R>, dim3(lanes), dim3(64), 0, b->stream, jobs, nm);
    BLAUNCHED(b, "k_step_end");
    return CNS_OK;
}

static int queue_batch_step(cns_batch *b) {
    for (const BatchGroup &g : b->groups) {
        const int rc = g.fp32 ? queue_group_step<float>(b, g) : queue_group_step<double>(b, g);
        if (rc) return rc;
    }
    return CNS_OK;
}

static void drop_graph(cns_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// capture one batch step as a graph (launch errors surface at capture end)
static int capture(cns_batch *b) {
    drop_graph(b);
    if (b->sync_check) return CNS_OK;  // checked mode launches eagerly
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = queue_batch_step(b);
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &gr);
    if (rc) {
        if (gr) (void)hipGraphDestroy(gr);
        return rc;
    }
    if (e != hipSuccess) return bfail(b, CNS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    b->graph = gr;
    BCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return CNS_OK;
}

extern "C" int cns_batch_create(cns_solver *const *members, int n, cns_batch **out) {
    if (!members || n <= 0 || !out) return bfail(nullptr, CNS_ERR_ARG, "cns_batch_create: bad arguments");
    *out = nullptr;
    for (int k = 0; k < n; k++) {
        cns_solver *m = members[k];
        if (!m) return bfail(nullptr, CNS_ERR_ARG, "cns_batch_create: null member");
        if (!m->have_state) return bfail(nullptr, CNS_ERR_STATE, "cns_batch_create: member without a state");
        if (m->in_batch) return bfail(nullptr, CNS_ERR_STATE, "cns_batch_create: member already in a batch");
        if (m->dev != members[0]->dev) return bfail(nullptr, CNS_ERR_ARG, "cns_batch_create: members on different devices");
        if (m->cfg.boundary_layer)
            return bfail(nullptr, CNS_ERR_ARG, "cns_batch_create: boundary-layer solvers run alone, not in batches");
        for (int q = 0; q < k; q++)
            if (members[q] == m) return bfail(nullptr, CNS_ERR_ARG, "cns_batch_create: repeated member");
    }
    cns_batch *b = new cns_batch();
    b->dev = members[0]->dev;
    {
        const char *sc = std::getenv("CNS_SYNC_CHECK");
        b->sync_check = sc && sc[0] == '1';
    }
    auto bail = [&](int code) {
        g_err = b->err;
        cns_batch_destroy(b);
        return code;
    };
    if (hipSetDevice(b->dev) != hipSuccess) {
        b->err = "hipSetDevice failed";
        return bail(CNS_ERR_HIP);
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, b->dev) == hipSuccess && prop.multiProcessorCount > 0)
        b->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        b->err = "hipStreamCreate failed";
        return bail(CNS_ERR_HIP);
    }
    for (int k = 0; k < n; k++) {
        cns_solver *m = members[k];
        if (hipStreamSynchronize(m->stream) != hipSuccess) {
            b->err = "member stream synchronisation failed";
            return bail(CNS_ERR_HIP);
        }
        b->members.push_back(m);
        b->cur0.push_back(m->cur);
        b->phase0.push_back(m->bsteps);
        // group key: kernel template and scheme
        const bool fp32 = m->fp32, wallx = !m->cfg.periodic_x, nonuni = m->cfg.nonuniform_x != 0,
                   unitv = m->cfg.visc_exp == 1.0;
        BatchGroup *g = nullptr;
        for (BatchGroup &x : b->groups)
            if (x.fp32 == fp32 && x.S == m->S && x.SV == m->SV && x.wallx == wallx && x.nonuni == nonuni &&
                x.unitv == unitv && x.scheme == m->cfg.scheme)
                g = &x;
        if (!g) {
            b->groups.emplace_back();
            g = &b->groups.back();
            g->fp32 = fp32;
            g->S = m->S;
            g->SV = m->SV;
            g->wallx = wallx;
            g->nonuni = nonuni;
            g->unitv = unitv;
            g->scheme = m->cfg.scheme;
            g->nst = n_stages(m);
            // untuned shape: one workgroup per item (the dispatcher balances
            // members of different sizes); short pencils for FP32 storage,
            // whose kernels run two workgroups per CU
            g->grid = 0;
            g->pencil = fp32 ? 16 : 32;
        }
        g->m.push_back(m);
        g->forcing = g->forcing || m->cfg.forcing_z;
        // step control: the count continues from the member's earlier batches
        StepCtl c{};
        c.istep = (int)m->bsteps;
        c.vfs = m->visc_from_state ? 1 : 0;
        c.check_cfl = m->cfg.check_cfl;
        c.forcing = m->cfg.forcing_z;
        c.cfl = m->cfg.cfl;
        c.vol = m->cfg.len_x * m->cfg.len_y * m->cfg.len_z;
        if (hipMemcpy(m->d_ctl, &c, sizeof(c), hipMemcpyHostToDevice) != hipSuccess) {
            b->err = "step control upload failed";
            return bail(CNS_ERR_HIP);
        }
    }
    for (BatchGroup &g : b->groups) {
        int rc = upload(b, g, g.pencil, false);
        if (rc) return bail(rc);
    }
    int rc = capture(b);
    if (rc) return bail(rc);
    for (cns_solver *m : b->members) m->in_batch = true;
    *out = b;
    return CNS_OK;
}

// Time of stage `st` of group g (dilatation + RHS, scratch outputs), best of `reps`.
template <typename R>
static int time_stage(cns_batch *b, const BatchGroup &g, int st, int reps, float &best) {
    hipEvent_t e0, e1;
    BCHK(b, hipEventCreate(&e0));
    BCHK(b, hipEventCreate(&e1));
    best = 1e30f;
    int rc = CNS_OK;
    for (int r = 0; r < reps && !rc; r++) {
        BCHK(b, hipEventRecord(e0, b->stream));
        rc = queue_stage<R>(b, g, st);
        if (rc) break;
        BCHK(b, hipEventRecord(e1, b->stream));
        BCHK(b, hipEventSynchronize(e1));
        float ms = 0.f;
        BCHK(b, hipEventElapsedTime(&ms, e0, e1));
        best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// Cost of a launch shape: best of five runs of the middle stage (after one
// unmeasured run that absorbs first-launch and cache effects).
static int time_stage_any(cns_batch *b, const BatchGroup &g, float &best) {
    const int st = g.nst / 2;
    float warm;
    int rc = g.fp32 ? time_stage<float>(b, g, st, 1, warm) : time_stage<double>(b, g, st, 1, warm);
    if (rc) return rc;
    return g.fp32 ? time_stage<float>(b, g, st, 5, best) : time_stage<double>(b, g, st, 5, best);
}

extern "C" int cns_batch_tune(cns_batch *b) {
    if (!b) return bfail(nullptr, CNS_ERR_ARG, "null batch");
    if (b->pending) {
        const int rc = cns_batch_wait(b);
        if (rc) return rc;
    }
    BCHK(b, hipSetDevice(b->dev));
    for (BatchGroup &g : b->groups) {
        int maxz = 0;
        for (cns_solver *m : g.m) maxz = std::max(maxz, m->cfg.nz);
        // 1. pencil length (z-chunk of a workgroup's march)
        std::vector<int> pencils;
        for (int L : {8, 16, 32, 64, 128, 256, 512})
            if (L >= 2 * g.S && L <= std::max(maxz, 2 * g.S)) pencils.push_back(L);
        if (pencils.empty()) pencils.push_back(maxz);
        int bestL = g.pencil;
        float bt = 1e30f;
        g.grid = 4 * b->cus;
        g.remap = 1;
        for (int L : pencils) {
            int rc = upload(b, g, L, true);
            if (rc) return rc;
            float t;
            if ((rc = time_stage_any(b, g, t))) return rc;
            if (t < bt) {
                bt = t;
                bestL = L;
            }
        }
        // 2. persistent workgroups (multiples of the CU count keep whole XCDs
        //    busy), jointly with the best pencil length, its half and double
        int bestG = g.grid;
        bt = 1e30f;
        int rc = CNS_OK;
        for (int L : pencils) {
            if (L != bestL && L != 2 * bestL && 2 * L != bestL) continue;
            if ((rc = upload(b, g, L, true))) return rc;
            for (int f : {1, 2, 4, 8, 0}) {
                g.grid = f * b->cus;  // 0: one workgroup per item
                float t;
                if ((rc = time_stage_any(b, g, t))) return rc;
                if (t < bt) {
                    bt = t;
                    bestG = g.grid;
                    g.pencil = L;
                }
            }
        }
        g.grid = bestG;
        if ((rc = upload(b, g, g.pencil, true))) return rc;
        // 3. XCD-contiguous item order on or off
        float t0, t1;
        g.remap = 0;
        if ((rc = time_stage_any(b, g, t0))) return rc;
        g.remap = 1;
        if ((rc = time_stage_any(b, g, t1))) return rc;
        g.remap = t1 <= t0 ? 1 : 0;
        if ((rc = upload(b, g, g.pencil, false))) return rc;
    }
    return capture(b);
}

static int batch_run(cns_batch *b, int nsteps) {
    for (int i = 0; i < nsteps; i++) {
        if (b->exec) {
            BCHK(b, hipGraphLaunch(b->exec, b->stream));
        } else {
            int rc = queue_batch_step(b);
            if (rc) return rc;
        }
    }
    return CNS_OK;
}

static void batch_after(cns_batch *b, int nsteps) {
    b->steps += nsteps;
    if (nsteps > 0)
        for (size_t k = 0; k < b->members.size(); k++) {
            cns_solver *m = b->members[k];
            m->visc_from_state = false;
            m->cur = (b->steps & 1) ? next_cur(m, b->cur0[k]) : b->cur0[k];
            m->bsteps = b->phase0[k] + b->steps;
        }
}

extern "C" int cns_batch_step(cns_batch *b, int nsteps) {
    if (!b) return bfail(nullptr, CNS_ERR_ARG, "null batch");
    if (b->pending) {
        const int rc = cns_batch_wait(b);
        if (rc) return rc;
    }
    if (nsteps < 0) return bfail(b, CNS_ERR_ARG, "cns_batch_step: nsteps < 0");
    BCHK(b, hipSetDevice(b->dev));
    int rc = batch_run(b, nsteps);
    if (rc) return rc;
    BCHK(b, hipStreamSynchronize(b->stream));
    batch_after(b, nsteps);
    return CNS_OK;
}

extern "C" int cns_batch_launch(cns_batch *b, int nsteps) {
    if (!b) return bfail(nullptr, CNS_ERR_ARG, "null batch");
    if (nsteps < 0) return bfail(b, CNS_ERR_ARG, "cns_batch_launch: nsteps < 0");
    BCHK(b, hipSetDevice(b->dev));
    int rc = batch_run(b, nsteps);
    if (rc) return rc;
    b->pending += nsteps;
    return CNS_OK;
}

extern "C" int cns_batch_wait(cns_batch *b) {
    if (!b) return bfail(nullptr, CNS_ERR_ARG, "null batch");
    BCHK(b, hipSetDevice(b->dev));
    BCHK(b, hipStreamSynchronize(b->stream));
    const int n = b->pending;
    b->pending = 0;
    batch_after(b, n);
    return CNS_OK;
}

extern "C" int cns_batch_time_steps(cns_batch *b, int nsteps, double *ms) {
    if (!b || !ms) return bfail(b, CNS_ERR_ARG, "cns_batch_time_steps: null argument");
    if (b->pending) {
        const int rc = cns_batch_wait(b);
        if (rc) return rc;
    }
    if (nsteps < 0) return bfail(b, CNS_ERR_ARG, "cns_batch_time_steps: nsteps < 0");
    BCHK(b, hipSetDevice(b->dev));
    hipEvent_t e0, e1;
    BCHK(b, hipEventCreate(&e0));
    BCHK(b, hipEventCreate(&e1));
    BCHK(b, hipEventRecord(e0, b->stream));
    int rc = batch_run(b, nsteps);
    if (!rc) {
        BCHK(b, hipEventRecord(e1, b->stream));
        BCHK(b, hipEventSynchronize(e1));
        float t = 0.f;
        BCHK(b, hipEventElapsedTime(&t, e0, e1));
        *ms = t;
        batch_after(b, nsteps);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

extern "C" int cns_batch_set_shape(cns_batch *b, int g, int pencil, int grid, int remap) {
    if (!b) return bfail(nullptr, CNS_ERR_ARG, "null batch");
    if (g < 0 || g >= (int)b->groups.size() || pencil < 1 || grid < 0)
        return bfail(b, CNS_ERR_ARG, "cns_batch_set_shape: bad group, pencil or grid");
    if (b->pending) {
        const int rc = cns_batch_wait(b);
        if (rc) return rc;
    }
    BCHK(b, hipSetDevice(b->dev));
    BatchGroup &x = b->groups[g];
    x.pencil = pencil;
    x.grid = grid;
    x.remap = remap ? 1 : 0;
    int rc = upload(b, x, x.pencil, false);
    if (rc) return rc;
    return capture(b);
}

extern "C" int cns_batch_groups(const cns_batch *b) { return b ? (int)b->groups.size() : 0; }

extern "C" int cns_batch_group_info(const cns_batch *b, int g, int *n_members, int *pencil, int *grid, int *remap,
                                    int *work_items) {
    if (!b || g < 0 || g >= (int)b->groups.size()) return CNS_ERR_ARG;
    const BatchGroup &x = b->groups[g];
    if (n_members) *n_members = (int)x.m.size();
    if (pencil) *pencil = x.pencil;
    if (grid) *grid = x.grid;
    if (remap) *remap = x.remap;
    if (work_items) *work_items = x.nwork;
    return CNS_OK;
}

extern "C" int cns_batch_destroy(cns_batch *b) {
    if (!b) return CNS_OK;
    (void)hipSetDevice(b->dev);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->pending) batch_after(b, b->pending);
    b->pending = 0;
    drop_graph(b);
    for (BatchGroup &g : b->groups) dev_free(b->dev, g.d_jobs);
    for (cns_solver *m : b->members) m->in_batch = false;
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return CNS_OK;
}

extern "C" const char *cns_batch_last_error(const cns_batch *b) { return b ? b->err.c_str() : g_err.c_str(); }
