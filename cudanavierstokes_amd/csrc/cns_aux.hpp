// cns_aux.hpp -- pointwise and reduction kernels around the RHS:
// state conversion (primitive <-> conserved), CFL time step, volume
// integrals for the forcing controller and bulk statistics, and the
// on-device step bookkeeping (dt, dp/dz, time) that replaces the
// reference's host round trips (src/cuda_main.cu:249-265).
#pragma once
#include "cns_device.hpp"

namespace cns {

constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;  // fixed partial count: deterministic sums
constexpr int NPART = 4;          // rho, rho*w, rho*E, u.u

// Device-resident scalars of one solver.
struct Scalars {
    double dt, dpdz, time, pad;
    unsigned long long conv_max, visc_max;
    double part_out[NPART];
};

// (rho, u, v, w, rhoE) -> (rho, rho u, rho v, rho w, rhoE), in place
template <typename R>
__global__ void k_to_conserved(long n, R *r, R *u, R *v, R *w) {
    for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
        const R rr = r[g];
        u[g] = u[g] * rr;
        v[g] = v[g] * rr;
        w[g] = w[g] * rr;
    }
}

// conserved -> primitive velocity into separate outputs
template <typename R>
__global__ void k_to_velocity(long n, const R *r, const R *ru, const R *rv, const R *rw, double *u, double *v,
                              double *w) {
    for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
        const R rr = r[g];
        u[g] = (double)(ru[g] / rr);
        v[g] = (double)(rv[g] / rr);
        w[g] = (double)(rw[g] / rr);
    }
}

__device__ __forceinline__ double dmax(double a, double b) { return a >= b ? a : b; }

// Block maximum of two non-negative values, then one atomic max each.
// Positive doubles order like their bit patterns, so an unsigned 64-bit
// atomic max is exact.
__device__ __forceinline__ void block_max2_to(double a, double b, unsigned long long *outa, unsigned long long *outb) {
    __shared__ double ra[RED_THREADS / 64], rb[RED_THREADS / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a = fmax(a, __shfl_xor(a, off));
        b = fmax(b, __shfl_xor(b, off));
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        ra[wid] = a;
        rb[wid] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ma = ra[0], mb = rb[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            ma = fmax(ma, ra[w]);
            mb = fmax(mb, rb[w]);
        }
        atomicMax(outa, (unsigned long long)__double_as_longlong(ma));
        if (outb) atomicMax(outb, (unsigned long long)__double_as_longlong(mb));
    }
}

// Local CFL inverse time scales (src/calc_stress.cu:140-160):
//   conv = max((|u|+c)/dx_i, (|v|+c)/dy, (|w|+c)/dz),  c = sqrt(g(g-1) e_int)
//   visc = max(mu/dx_i^2, mu/dy^2, mu/dz^2)   (only with visc_from_state; the
//   solver loop otherwise takes mu from the last RHS input, as the reference
//   reads the mu field of the last calcState)
template <typename R>
__device__ __forceinline__ void cfl_point(const Params &P, const StagePtrs<R> &A, long g, int visc_from_state,
                                          double &cmax, double &vmax) {
    const int i = (int)(g % P.mx);
    const long fs = A.fs;
    // full values (the offsets are zero for FP64 storage)
    const double r = (double)A.qin[g] + P.rho0;
    const double u = (double)A.qin[g + fs] / r, v = (double)A.qin[g + 2 * fs] / r, w = (double)A.qin[g + 3 * fs] / r;
    const double ien = ((double)A.qin[g + 4 * fs] + P.E0) / r - 0.5 * (u * u + v * v + w * w);
    const double sos = sqrt(P.gam * (P.gam - 1) * ien);
    const double dx = P.dxv[i];
    const double c1 = (fabs(u) + sos) / dx, c2 = (fabs(v) + sos) * P.d_dy, c3 = (fabs(w) + sos) * P.d_dz;
    cmax = dmax(cmax, dmax(c1, dmax(c2, c3)));
    if (visc_from_state) {
        const double t = P.cvInv * ien;
        const double mu = (P.unit_visc ? t : visc_pow(t, P.viscexp)) / P.Re;
        const double v1 = mu / (dx * dx), v2 = mu * P.d_d2y, v3 = mu * P.d_d2z;
        vmax = dmax(vmax, dmax(v1, dmax(v2, v3)));
    }
}

template <typename R>
__global__ void k_cfl(Params P, StagePtrs<R> A, int visc_from_state, Scalars *sc) {
    const long mxy = (long)P.mx * P.my, n = mxy * P.mz;
    double cmax = 0.0, vmax = 0.0;
    for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
        cfl_point(P, A, g, visc_from_state, cmax, vmax);
    block_max2_to(cmax, vmax, &sc->conv_max, visc_from_state ? &sc->visc_max : nullptr);
}

// dt = min(1e5, CFL / max(conv, visc)), which equals the reference's
// min over points of CFL / max(conv_i, visc_i) (division is monotone).
__device__ __forceinline__ void dt_finalize(double cfl, Scalars *sc) {
    const double cm = __longlong_as_double((long long)sc->conv_max);
    const double vm = __longlong_as_double((long long)sc->visc_max);
    const double m = cm >= vm ? cm : vm;
    const double val = cfl / m;
    sc->dt = (100000.0 <= val) ? 100000.0 : val;
    sc->conv_max = 0ull;
}

static __global__ void k_dt_finalize(double cfl, Scalars *sc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) dt_finalize(cfl, sc);
}

static __global__ void k_reset_visc(Scalars *sc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) sc->visc_max = 0ull;
}

// time[istep] = time + dt (src/calc_stress.cu:12-14, src/cuda_math.cu:17-19)
static __global__ void k_advance_time(Scalars *sc, double *time_out, int istep) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc->time = sc->time + sc->dt;
        if (time_out) time_out[istep] = sc->time;
    }
}

// Per-block partial volume integrals, weight dxv[i]/d_dy/d_dz
// (src/cuda_math.cu:143-171): rho, rho*w, rho*E, u.u.
template <typename R>
__device__ __forceinline__ void partials_block(const Params &P, const StagePtrs<R> &A, double *part) {
    const long mxy = (long)P.mx * P.my, n = mxy * P.mz;
    double s[NPART] = {0, 0, 0, 0};
    for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
        const int i = (int)(g % P.mx);
        const double wgt = P.dxv[i];
        const long fs = A.fs;
        const double r = (double)A.qin[g] + P.rho0;
        const double ru = (double)A.qin[g + fs], rv = (double)A.qin[g + 2 * fs], rw = (double)A.qin[g + 3 * fs];
        const double u = ru / r, v = rv / r, w = rw / r;
        s[0] += r * wgt / P.d_dy / P.d_dz;
        s[1] += rw * wgt / P.d_dy / P.d_dz;
        s[2] += ((double)A.qin[g + 4 * fs] + P.E0) * wgt / P.d_dy / P.d_dz;
        s[3] += (u * u + v * v + w * w) * wgt / P.d_dy / P.d_dz;
    }
    __shared__ double sh[NPART][RED_THREADS];
    for (int c = 0; c < NPART; c++) sh[c][threadIdx.x] = s[c];
    __syncthreads();
    for (int size = RED_THREADS / 2; size > 0; size /= 2) {
        if ((int)threadIdx.x < size)
            for (int c = 0; c < NPART; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + size];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < NPART; c++) part[c * RED_BLOCKS + blockIdx.x] = sh[c][0];
}

template <typename R>
__global__ void k_partials(Params P, StagePtrs<R> A, double *part) {
    partials_block(P, A, part);
}

// Final fixed-order sum of the partials, then either the pressure-gradient
// controller (src/calc_stress.cu:16-18,98-120) or the bulk statistics
// (src/calc_stress.cu:162-201).  what: 1 controller, 2 bulk.
__device__ __forceinline__ void final_sums(const double *part, Scalars *sc, int what, int forcing, double vol,
                                           double *par1, double *par2, int istep) {
    __shared__ double sh[NPART][RED_THREADS];
    for (int c = 0; c < NPART; c++) {
        double a = 0.0;
        for (int b = threadIdx.x; b < RED_BLOCKS; b += RED_THREADS) a += part[c * RED_BLOCKS + b];
        sh[c][threadIdx.x] = a;
    }
    __syncthreads();
    for (int size = RED_THREADS / 2; size > 0; size /= 2) {
        if ((int)threadIdx.x < size)
            for (int c = 0; c < NPART; c++) sh[c][threadIdx.x] += sh[c][threadIdx.x + size];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int c = 0; c < NPART; c++) sc->part_out[c] = sh[c][0];
        const double rbulk = sh[0][0], rwint = sh[1][0], eint = sh[2][0], uu = sh[3][0];
        if (what == 1) {
            const double prev = sc->dpdz;
            sc->dpdz = 0.99 * prev - 0.5 * (rwint / rbulk - 1);
        } else if (what == 2) {
            if (forcing) {
                if (par1) par1[istep] = rwint / rbulk;
                if (par2) par2[istep] = eint;
            } else {
                if (par1) par1[istep] = uu / vol;
                if (par2) par2[istep] = __longlong_as_double(0x7ff8000000000000ll);
            }
        }
    }
}

static __global__ void k_final(const double *part, Scalars *sc, int what, int forcing, double vol, double *par1,
                        double *par2, int istep) {
    final_sums(part, sc, what, forcing, vol, par1, par2, istep);
}

static __global__ void k_fill(double *a, long n, double v) {
    for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) a[g] = v;
}

// Initial conditions evaluated on the device: the closed forms of the host
// fills (Taylor-Green vortex of src/init.cpp:126-148; the seeded channel),
// written straight into a state set.
struct InitArgs {
    int kind;  // 0 Taylor-Green, 1 channel
    int mx, my, mz;
    const double *x, *y, *z;
    double lx, ly, lz, Rgas, gam, ubulk, amp;
    unsigned long long seed;
};

__device__ __forceinline__ double splitmix_uniform(unsigned long long state) {
    unsigned long long z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// conserved (rho, rho u, rho v, rho w, rho E) of point g
__device__ __forceinline__ void init_point(const InitArgs &a, long g, double q[5]) {
    const int i = (int)(g % a.mx);
    double r, u, v, w, e;
    if (a.kind == 0) {
        const long jk = g / a.mx;
        const int j = (int)(jk % a.my), k = (int)(jk / a.my);
        const double fx = 2 * M_PI * a.x[i] / a.lx, fy = 2 * M_PI * a.y[j] / a.ly, fz = 2 * M_PI * a.z[k] / a.lz;
        u = sin(fx) * cos(fy) * cos(fz);
        v = -cos(fx) * sin(fy) * cos(fz);
        w = 0.0;
        const double press = a.Rgas + 1.0 / 16.0 * (cos(2.0 * fx) + cos(2.0 * fy)) * (cos(2.0 * fz) + 2.0);
        r = press / a.Rgas;
        e = press / (a.gam - 1.0) + 0.5 * r * (u * u + v * v + w * w);
    } else {
        // draws 3g, 3g+1, 3g+2 of the SplitMix64 stream: states seed + (3g+1+d) * increment
        const unsigned long long inc = 0x9E3779B97F4A7C15ull, base = a.seed + (unsigned long long)(3 * g) * inc;
        const double a1 = splitmix_uniform(base + inc), a2 = splitmix_uniform(base + 2 * inc),
                     a3 = splitmix_uniform(base + 3 * inc);
        const double xi = a.x[i] / (a.lx / 2.0) - 1.0;
        const double wm = 1.5 * a.ubulk * (1.0 - xi * xi);
        u = a.amp * a.ubulk * (2.0 * a1 - 1.0);
        v = a.amp * a.ubulk * (2.0 * a2 - 1.0);
        w = wm + a.amp * a.ubulk * (2.0 * a3 - 1.0);
        r = 1.0;
        e = a.Rgas / (a.gam - 1.0) + 0.5 * r * (u * u + v * v + w * w);
    }
    q[0] = r;
    q[1] = u * r;
    q[2] = v * r;
    q[3] = w * r;
    q[4] = e;
}

// per-block sums of rho and rho E (fixed RED_BLOCKS partials), for the FP32
// reference state
static __global__ void k_init_sums(InitArgs a, long n, double *part) {
    double sr = 0.0, se = 0.0;
    for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
        double q[5];
        init_point(a, g, q);
        sr += q[0];
        se += q[4];
    }
    __shared__ double sh[2][RED_THREADS];
    sh[0][threadIdx.x] = sr;
    sh[1][threadIdx.x] = se;
    __syncthreads();
    for (int size = RED_THREADS / 2; size > 0; size /= 2) {
        if ((int)threadIdx.x < size) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + size];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + size];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sh[0][0];
        part[RED_BLOCKS + blockIdx.x] = sh[1][0];
    }
}

// store the state: FP64 as is; FP32 as offsets from (rho0, E0) split into
// hi = fl(x) and lo = fl(x - hi) (lo may be null)
template <typename R>
__global__ void k_init_store(InitArgs a, long n, R *set, R *lo, double rho0, double E0) {
    for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x) {
        double q[5];
        init_point(a, g, q);
        q[0] -= rho0;
        q[4] -= E0;
#pragma unroll
        for (int c = 0; c < 5; c++) {
            const R h = (R)q[c];
            set[g + c * n] = h;
            if (lo) lo[g + c * n] = (R)(q[c] - (double)h);
        }
    }
}

}  // namespace cns
