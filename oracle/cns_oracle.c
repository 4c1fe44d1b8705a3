/*
 * cns_oracle.c -- CPU FP64 restatement of the reference solver's explicit
 * time-advance path (simone-silvestri/CudaNavierStokes).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product package
 * links, imports or calls it.
 *
 * Structure follows the reference call stack one function per reference
 * routine, with plain loops over pencils:
 *   calc_state      <- src/cuda_main.cu:218-242            (calcState)
 *   der_vel         <- src/calc_stress.cu:20-63            (derVelX/Y/Z)
 *                      src/cuda_derivs.h:192-205,314-386
 *   dil             <- src/calc_stress.cu:87-96            (calcDil)
 *   rhs_x           <- src/cuda_rhs.cu:9-128               (deviceRHSX, writes =)
 *   rhs_y           <- src/cuda_rhs.cu:130-262             (deviceRHSY, adds +=)
 *   rhs_z           <- src/cuda_rhs.cu:264-396             (deviceRHSZ, adds +=, forcing)
 *   add_sponge      <- src/sponge.cu:31-41                 (addSponge)
 *   ghost fills     <- src/boundary.h:38-180, src/boundary_condition_{x,y,z}.h
 *                      (periodic, isothermal wall, boundary layer)
 *   blowing         <- src/perturbation.h:25-53                (PerturbUvel)
 *   stencils/fluxes <- src/cuda_derivs.h:30-312
 *   SMR RK3         <- src/cuda_main.cu:10-11,109-186,244-247 (race-free: the
 *                      momentum is formed with the density before the update)
 *   Kutta RK3       <- src/cuda_main.cu:44-107,188-216
 *   classic RK4     <- no reference counterpart; closed form of the method
 *   calc_dt         <- src/calc_stress.cu:122-160, src/cuda_math.cu:232-259
 *   forcing         <- src/calc_stress.cu:16-18,98-120, src/cuda_utils.cu:68-70
 *   bulk            <- src/calc_stress.cu:162-201, src/cuda_math.cu:143-230
 *   grid/metrics    <- src/init.cpp:32-92,258-277, src/cuda_utils.cu:60-139
 *   coefficients    <- src/globals.h:69-95 (computed here from the closed form)
 *
 * Parity pins (tests/test_oracle_*.py): coefficient tables, modified
 * wavenumbers, polynomial exactness, split-flux identities, free stream,
 * discrete conservation, manufactured-solution convergence of the full RHS
 * against a symbolic continuous RHS (periodic and isothermal-wall cases),
 * RK stability polynomials, dt and reduction closed forms, boundary-layer
 * ghost rules (exact derivatives of linear fields at the extrapolated ends,
 * adiabatic rest state, pressure extrapolated on its own) and the blowing's
 * band, zeros, antisymmetry, spanwise cosine and hand-evaluated values.
 * Unpinned: libm pow/tanh last-ulp behaviour against CUDA libdevice.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OMAX(a, b) ((a) >= (b) ? (a) : (b))
#define OMIN(a, b) ((a) <= (b) ? (a) : (b))

typedef struct orc_config {
    int mx, my, mz;
    int s, sv;          /* advective / viscous stencil half widths */
    int periodic_x;     /* 1: periodic in x; 0: isothermal no-slip walls */
    int nonuniform_x;   /* 1: tanh-stretched x with metric terms */
    int forcing;        /* 1: pressure-gradient controller in z */
    int sponge;         /* 1: add sponge relaxation */
    int scheme;         /* 0 SMR low-storage RK3, 1 Kutta RK3, 2 classic RK4 */
    int check_cfl;
    int check_bulk;
    int fix_y_work;     /* 0: reference viscous-work term in y; 1: corrected */
    int boundary_layer; /* 1 (with periodic_x 0): boundary-layer ghost rules in x and z */
    int perturbed;      /* 1: wall-normal blowing at the bottom x boundary (PerturbUvel) */
    double Lx, Ly, Lz;
    double Re, Pr, Ma, gam, viscexp, cfl;
    double stretch, twall_bot, twall_top;
    double dpdz0;
} orc_config;

typedef struct orc {
    orc_config c;
    long n;
    int S, SV;
    double *x, *y, *z, *xp, *xpp, *dxv;
    double dx, d_dx, d_dy, d_dz, d_d2x, d_d2y, d_d2z;
    double cF[4], cS[5], cVF[4], cVS[5];
    double *VSx; /* (2*SV+1) x mx stretched second-derivative weights */
    double Rgas, Ec;
    double *r, *u, *v, *w, *e;          /* state: rho, u, v, w, rhoE */
    double *h, *t, *p, *m, *l;          /* calcState outputs */
    double *g[9];                       /* velocity gradients */
    double *dil;
    double *R1[5], *R2[5], *R3[5], *O[5];
    double dt, dpdz, time;
    double *spx, *spz, *ref[5];
} orc;

/* ------------------------------------------------------------------ */
/* coefficients: exact rationals of the central 2s-order weights       */
/* ------------------------------------------------------------------ */
static int64_t gcd64(int64_t a, int64_t b) {
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) { int64_t t = a % b; a = b; b = t; }
    return a ? a : 1;
}
static int64_t fact(int n) { int64_t f = 1; for (int i = 2; i <= n; i++) f *= i; return f; }

/* cF[it] multiplies f[i+it-s]-f[i+s-it]; cS[it] (it<s) multiplies
 * f[i+it-s]+f[i+s-it]; cS[s] is the centre weight (reference layout,
 * src/globals.h:69-95). */
void orc_coeffs(int s, double *cF, double *cS) {
    int64_t sf2 = fact(s) * fact(s);
    int64_t n0 = 0, d0 = 1; /* running sum of second-derivative weights */
    for (int k = 1; k <= s; k++) {
        int64_t den = (int64_t)k * fact(s - k) * fact(s + k);
        int64_t num = sf2;
        int64_t g = gcd64(num, den);
        num /= g; den /= g;
        double sign = (k % 2 == 1) ? 1.0 : -1.0; /* (-1)^(k+1) */
        /* first derivative weight on f[i+k]: sign*num/den; coeffF[s-k] = -that */
        cF[s - k] = (-sign * (double)num) / (double)den;
        /* second derivative weight on f[i+-k]: 2*sign*num/(den*k) */
        int64_t n2 = 2 * num, dd2 = den * k;
        int64_t g2 = gcd64(n2, dd2);
        n2 /= g2; dd2 /= g2;
        cS[s - k] = (sign * (double)n2) / (double)dd2;
        /* accumulate sign*n2/dd2 into n0/d0 exactly */
        int64_t sn = (k % 2 == 1) ? n2 : -n2;
        int64_t nn = n0 * dd2 + sn * d0, dn = d0 * dd2;
        int64_t gg = gcd64(nn, dn);
        n0 = nn / gg; d0 = dn / gg;
    }
    /* centre weight = -2 * sum */
    int64_t nc = -2 * n0, dc = d0;
    int64_t gc = gcd64(nc, dc);
    nc /= gc; dc /= gc;
    cS[s] = (double)nc / (double)dc;
}

/* ------------------------------------------------------------------ */
/* grid (src/init.cpp:32-92, 258-277; src/cuda_utils.cu:60-139)         */
/* ------------------------------------------------------------------ */
static void deriv_grid(const orc *o, double *d2f, double *df, const double *f, double dx) {
    int mx = o->c.mx, S = o->S;
    double *fb = (double *)malloc(sizeof(double) * (mx + 2 * S));
    for (int i = S; i < mx + S; i++) fb[i] = f[i - S];
    for (int i = 0; i < S; i++) {
        fb[i] = -fb[2 * S - i - 1];
        fb[mx + S + i] = 2 * o->c.Lx - fb[mx + S - i - 1];
    }
    for (int i = 0; i < mx; i++) {
        df[i] = 0.0;
        d2f[i] = o->cS[S] * fb[i + S] / dx / dx;
        for (int it = 0; it < S; it++) {
            df[i] += o->cF[it] * (fb[i + it] - fb[i + S * 2 - it]) / dx;
            d2f[i] += o->cS[it] * (fb[i + it] + fb[i + S * 2 - it]) / dx / dx;
        }
    }
    free(fb);
}

static void init_grid(orc *o) {
    const orc_config *c = &o->c;
    int mx = c->mx, my = c->my, mz = c->mz;
    o->dx = c->Lx * (1.0) / (mx);
    /* a boundary layer takes the lower half of the tanh profile (clustered at
       the wall x = 0 only): src/init.cpp:41-50 */
    int denom = c->boundary_layer ? 2 * mx : mx;
    double denom2 = c->boundary_layer ? 1.0 : 2.0;
    double *xn = (double *)malloc(sizeof(double) * (mx + 1));
    for (int i = 0; i < mx + 1; i++)
        xn[i] = tanh(c->stretch * ((i * 1.0) / denom - 0.5)) / tanh(c->stretch * 0.5);
    for (int i = 0; i < mx; i++) o->x[i] = c->Lx * (1.0 + (xn[i] + xn[i + 1]) / 2.0) / denom2;
    free(xn);
    deriv_grid(o, o->xpp, o->xp, o->x, o->dx);
    for (int i = 0; i < mx; i++) o->xp[i] = 1.0 / o->xp[i];
    if (!c->nonuniform_x) {
        for (int i = 0; i < mx; i++) o->x[i] = c->Lx * (0.5 + i * 1.0) / (mx);
        o->dx = o->x[1] - o->x[0];
    }
    for (int j = 0; j < my; j++) o->y[j] = c->Ly * (0.5 + j * 1.0) / (my);
    for (int k = 0; k < mz; k++) o->z[k] = c->Lz * (0.5 + k * 1.0) / (mz);

    o->d_dx = 1.0 / (o->dx);
    o->d_dy = 1.0 / (o->y[1] - o->y[0]);
    o->d_dz = 1.0 / (o->z[1] - o->z[0]);
    o->d_d2x = o->d_dx * o->d_dx;
    o->d_d2y = o->d_dy * o->d_dy;
    o->d_d2z = o->d_dz * o->d_dz;

    o->dxv[0] = (o->x[1] + o->x[0]) / 2.0;
    for (int i = 1; i < mx - 1; i++) o->dxv[i] = (o->x[i + 1] - o->x[i - 1]) / 2.0;
    o->dxv[mx - 1] = c->Lx - (o->x[mx - 1] + o->x[mx - 2]) / 2.0;

    int SV = o->SV;
    const double *xp = o->xp, *xpp = o->xpp;
    for (int it = 0; it < SV; it++)
        for (int i = 0; i < mx; i++)
            o->VSx[i + it * mx] = (o->cVS[it] * (xp[i] * xp[i]) * o->d_d2x -
                                   o->cVF[it] * xpp[i] * (xp[i] * xp[i] * xp[i]) * o->d_dx);
    for (int i = 0; i < mx; i++) o->VSx[i + SV * mx] = o->cVS[SV] * (xp[i] * xp[i]) * o->d_d2x;
    for (int it = SV + 1; it < 2 * SV + 1; it++)
        for (int i = 0; i < mx; i++)
            o->VSx[i + it * mx] = (o->cVS[2 * SV - it] * (xp[i] * xp[i]) * o->d_d2x +
                                   o->cVF[2 * SV - it] * xpp[i] * (xp[i] * xp[i] * xp[i]) * o->d_dx);
}

/* ------------------------------------------------------------------ */
/* allocation                                                          */
/* ------------------------------------------------------------------ */
static double *fal(long n) { double *p = (double *)calloc((size_t)n, sizeof(double)); return p; }

int orc_config_size(void) { return (int)sizeof(orc_config); }

orc *orc_create(const orc_config *cfg) {
    if (cfg->s < 1 || cfg->s > 4 || cfg->sv < 1 || cfg->sv > cfg->s) return NULL;
    if (cfg->mx < 2 * cfg->s || cfg->my < 2 * cfg->s || cfg->mz < 2 * cfg->s) return NULL;
    if (cfg->boundary_layer && cfg->periodic_x) return NULL;
    orc *o = (orc *)calloc(1, sizeof(orc));
    o->c = *cfg;
    o->S = cfg->s;
    o->SV = cfg->sv;
    o->n = (long)cfg->mx * cfg->my * cfg->mz;
    orc_coeffs(o->S, o->cF, o->cS);
    orc_coeffs(o->SV, o->cVF, o->cVS);
    o->Ec = ((cfg->gam - 1.0) * cfg->Ma * cfg->Ma);
    o->Rgas = (1.0 / (cfg->gam * cfg->Ma * cfg->Ma));
    o->x = fal(cfg->mx); o->xp = fal(cfg->mx); o->xpp = fal(cfg->mx); o->dxv = fal(cfg->mx);
    o->y = fal(cfg->my); o->z = fal(cfg->mz);
    o->VSx = fal((long)cfg->mx * (2 * o->SV + 1));
    init_grid(o);
    long n = o->n;
    o->r = fal(n); o->u = fal(n); o->v = fal(n); o->w = fal(n); o->e = fal(n);
    o->h = fal(n); o->t = fal(n); o->p = fal(n); o->m = fal(n); o->l = fal(n);
    for (int q = 0; q < 9; q++) o->g[q] = fal(n);
    o->dil = fal(n);
    for (int q = 0; q < 5; q++) { o->R1[q] = fal(n); o->R2[q] = fal(n); o->R3[q] = fal(n); o->O[q] = fal(n); }
    o->spx = fal(cfg->mx); o->spz = fal(cfg->mz);
    for (int q = 0; q < 5; q++) o->ref[q] = fal((long)cfg->mx * cfg->mz);
    o->dpdz = cfg->forcing ? cfg->dpdz0 : 0.0;
    o->dt = 0.0;
    o->time = 0.0;
    return o;
}

void orc_destroy(orc *o) {
    if (!o) return;
    free(o->x); free(o->xp); free(o->xpp); free(o->dxv); free(o->y); free(o->z); free(o->VSx);
    free(o->r); free(o->u); free(o->v); free(o->w); free(o->e);
    free(o->h); free(o->t); free(o->p); free(o->m); free(o->l);
    for (int q = 0; q < 9; q++) free(o->g[q]);
    free(o->dil);
    for (int q = 0; q < 5; q++) { free(o->R1[q]); free(o->R2[q]); free(o->R3[q]); free(o->O[q]); free(o->ref[q]); }
    free(o->spx); free(o->spz);
    free(o);
}

/* ------------------------------------------------------------------ */
/* equation of state (src/cuda_main.cu:218-242)                        */
/* ------------------------------------------------------------------ */
static void calc_state(orc *o) {
    const orc_config *c = &o->c;
    double Rgas = o->Rgas, gam = c->gam;
    double cvInv = (gam - 1.0) / Rgas;
    long n = o->n;
#pragma omp parallel for schedule(static)
    for (long gl = 0; gl < n; gl++) {
        double invrho = 1.0 / o->r[gl];
        double en = o->e[gl] * invrho - 0.5 * (o->u[gl] * o->u[gl] + o->v[gl] * o->v[gl] + o->w[gl] * o->w[gl]);
        o->t[gl] = cvInv * en;
        o->p[gl] = o->r[gl] * Rgas * o->t[gl];
        o->h[gl] = (o->e[gl] + o->p[gl]) * invrho;
        double suth = pow(o->t[gl], c->viscexp);
        o->m[gl] = suth / c->Re;
        o->l[gl] = suth / c->Re / c->Pr / o->Ec;
    }
}

/* ------------------------------------------------------------------ */
/* pencil stencils (src/cuda_derivs.h)                                 */
/* b points to a buffer whose element b[si] is the pencil point;       */
/* halo of S (advective) cells on each side.                           */
/* ------------------------------------------------------------------ */
/* x-direction: sum, then * d_dx, then * xp (cuda_derivs.h:157-224) */
static inline double d1x(const orc *o, const double *b, int si, int i) {
    int S = o->S;
    double df = 0.0;
    for (int it = 0; it < S; it++) df += o->cF[it] * (b[si + it - S] - b[si + S - it]);
    df = df * o->d_dx;
    if (o->c.nonuniform_x) df = df * o->xp[i];
    return df;
}
static inline double dv1x(const orc *o, const double *b, int si, int i) {
    int SV = o->SV;
    double df = 0.0;
    for (int it = 0; it < SV; it++) df += o->cVF[it] * (b[si + it - SV] - b[si + SV - it]);
    df = df * o->d_dx;
    if (o->c.nonuniform_x) df = df * o->xp[i];
    return df;
}
static inline double dv2x(const orc *o, const double *b, int si, int i) {
    int SV = o->SV;
    double d2f;
    if (o->c.nonuniform_x) {
        d2f = 0.0;
        for (int it = 0; it < 2 * SV + 1; it++) d2f += o->VSx[it * o->c.mx + i] * (b[si + it - SV]);
    } else {
        d2f = o->cVS[SV] * b[si] * o->d_d2x;
        for (int it = 0; it < SV; it++) d2f += o->cVS[it] * (b[si + it - SV] + b[si + SV - it]) * o->d_d2x;
    }
    return d2f;
}
static inline double fquad_x(const orc *o, const double *f, const double *g, int si, int i) {
    int S = o->S;
    double flxp = 0.0, flxm = 0.0;
    for (int lt = 1; lt < S + 1; lt++)
        for (int mt = 0; mt < lt; mt++) {
            flxp -= o->cF[S - lt] * (f[si - mt] + f[si - mt + lt]) * (g[si - mt] + g[si - mt + lt]);
            flxm -= o->cF[S - lt] * (f[si - mt - 1] + f[si - mt + lt - 1]) * (g[si - mt - 1] + g[si - mt + lt - 1]);
        }
    double df = 0.5 * o->d_dx * (flxm - flxp);
    if (o->c.nonuniform_x) df = df * o->xp[i];
    return df;
}
static inline double fcube_x(const orc *o, const double *f, const double *g, const double *hh, int si, int i) {
    int S = o->S;
    double flxp = 0.0, flxm = 0.0;
    for (int lt = 1; lt < S + 1; lt++)
        for (int mt = 0; mt < lt; mt++) {
            flxp -= o->cF[S - lt] * (f[si - mt] + f[si - mt + lt]) * (g[si - mt] + g[si - mt + lt]) *
                    (hh[si - mt] + hh[si - mt + lt]);
            flxm -= o->cF[S - lt] * (f[si - mt - 1] + f[si - mt + lt - 1]) * (g[si - mt - 1] + g[si - mt + lt - 1]) *
                    (hh[si - mt - 1] + hh[si - mt + lt - 1]);
        }
    double df = 0.25 * o->d_dx * (flxm - flxp);
    if (o->c.nonuniform_x) df = df * o->xp[i];
    return df;
}
/* y/z-direction: spacing multiplied inside the sum (cuda_derivs.h:226-312) */
static inline double d1g(const double *cF, int S, double d, const double *b, int si) {
    double df = 0.0;
    for (int it = 0; it < S; it++) df += cF[it] * (b[si + it - S] - b[si + S - it]) * d;
    return df;
}
static inline double d2g(const double *cS, int S, double d2, const double *b, int si) {
    double d2f = cS[S] * b[si] * d2;
    for (int it = 0; it < S; it++) d2f += cS[it] * (b[si + it - S] + b[si + S - it]) * d2;
    return d2f;
}
static inline double fquad_g(const double *cF, int S, double d, const double *f, const double *g, int si) {
    double flxp = 0.0, flxm = 0.0;
    for (int lt = 1; lt < S + 1; lt++)
        for (int mt = 0; mt < lt; mt++) {
            flxp -= cF[S - lt] * (f[si - mt] + f[si - mt + lt]) * (g[si - mt] + g[si - mt + lt]);
            flxm -= cF[S - lt] * (f[si - mt - 1] + f[si - mt + lt - 1]) * (g[si - mt - 1] + g[si - mt + lt - 1]);
        }
    return 0.5 * d * (flxm - flxp);
}
static inline double fcube_g(const double *cF, int S, double d, const double *f, const double *g, const double *hh,
                             int si) {
    double flxp = 0.0, flxm = 0.0;
    for (int lt = 1; lt < S + 1; lt++)
        for (int mt = 0; mt < lt; mt++) {
            flxp -= cF[S - lt] * (f[si - mt] + f[si - mt + lt]) * (g[si - mt] + g[si - mt + lt]) *
                    (hh[si - mt] + hh[si - mt + lt]);
            flxm -= cF[S - lt] * (f[si - mt - 1] + f[si - mt + lt - 1]) * (g[si - mt - 1] + g[si - mt + lt - 1]) *
                    (hh[si - mt - 1] + hh[si - mt + lt - 1]);
        }
    return 0.25 * d * (flxm - flxp);
}

/* ------------------------------------------------------------------ */
/* ghost fills on a pencil buffer b[0 .. n+2S), interior at [S, S+n)   */
/* (src/boundary.h:38-148)                                             */
/* ------------------------------------------------------------------ */
static inline void g_per(double *b, int n, int S) {
    for (int l = 0; l < S; l++) {
        b[l] = b[l + n];
        b[l + S + n] = b[l + S];
    }
}
static inline void g_vel(double *b, int n, int S) { /* wallBCxVel: antisymmetric */
    for (int l = 0; l < S; l++) {
        b[l] = -b[2 * S - l - 1];
        b[l + S + n] = -b[n + S - l - 1];
    }
}
static inline void g_mir(double *b, int n, int S) { /* wallBCxMir: symmetric */
    for (int l = 0; l < S; l++) {
        b[l] = b[2 * S - l - 1];
        b[l + S + n] = b[n + S - l - 1];
    }
}
static inline void g_ext(double *b, int n, int S, double top, double bot) { /* wallBCxExt */
    for (int l = 0; l < S; l++) {
        b[l] = 2.0 * bot - b[2 * S - l - 1];
        b[l + S + n] = 2.0 * top - b[n + S - l - 1];
    }
}

/* boundary layer (src/boundary.h:150-180): linear extrapolation about the
   last node (topBCxExt, topBCzExt) and about the first node (botBCzExt);
   at the bottom x boundary the odd and even reflections about the face are
   g_vel's and g_mir's bottom halves (botBCxExt with value 0, botBCxMir) */
static inline void g_ext_top(double *b, int n, int S) {
    for (int d = 1; d <= S; d++) b[S + n - 1 + d] = 2.0 * b[S + n - 1] - b[S + n - 1 - d];
}
static inline void g_ext_bot(double *b, int S) {
    for (int d = 1; d <= S; d++) b[S - d] = 2.0 * b[S] - b[S + d];
}
static inline void g_odd_bot(double *b, int S) {
    for (int l = 0; l < S; l++) b[l] = -b[2 * S - l - 1];
}
static inline void g_mir_bot(double *b, int S) {
    for (int l = 0; l < S; l++) b[l] = b[2 * S - l - 1];
}

/* Wall-normal blowing of src/perturbation.h:25-53 (PerturbUvel): for global
   z index k in [kC - LP/2, kC + LP/2] every bottom ghost of u takes
   kappa g(xi) [amp1 sin(omega1 t) + amp2 sin(omega2 t) cos(y/lambda)] with
   g = 15.1875 xi^5 - 35.4375 xi^4 + 20.25 xi^3, kappa = +1 before kC and -1
   from kC on, xi the integer distance from the nearer end over LP/2,
   y = j/d_dy, lambda = Ly/(2 pi), omega1 = Re 121e-6, omega2 = 10, t the
   simulation time after the step's advance.  Returns 1 and the value when
   k is in the band. */
int orc_perturbation(const orc *o, int j, int k, double t, double *val) {
    const int kC = 110, LP = 40;
    const int kSt = kC - LP / 2, kEn = kC + LP / 2;
    if (k < kSt || k > kEn) return 0;
    int kappa, alpha, beta;
    if (k < kC) {
        kappa = 1;
        alpha = k - kSt;
        beta = kC - kSt;
    } else {
        kappa = -1;
        alpha = kEn - k;
        beta = kEn - kC;
    }
    double xi = alpha * 1.0 / beta;
    double g = (15.1875 * xi * xi * xi * xi * xi) - (35.4375 * xi * xi * xi * xi) + (20.25 * xi * xi * xi);
    double lam = o->c.Ly / (2.0 * M_PI);
    double om1 = o->c.Re * 121.e-6, om2 = 10.0;
    double y = j / o->d_dy;
    *val = 2e-4 * kappa * g * sin(om1 * t) + 3e-5 * kappa * g * sin(om2 * t) * cos(y / lam);
    return 1;
}

/* u's bottom ghosts at (j, k) when the blowing applies */
static inline void g_perturb(const orc *o, double *b, int S, int j, int k) {
    double val;
    if (o->c.boundary_layer && o->c.perturbed && orc_perturbation(o, j, k, o->time, &val))
        for (int l = 0; l < S; l++) b[l] = val;
}

/* ------------------------------------------------------------------ */
/* velocity gradients (src/calc_stress.cu:20-63)                       */
/* ------------------------------------------------------------------ */
static void der_vel(orc *o) {
    const orc_config *c = &o->c;
    int mx = c->mx, my = c->my, mz = c->mz, S = o->S, SV = o->SV;
    long mxy = (long)mx * my;
    const double *fld[3] = {o->u, o->v, o->w};
#pragma omp parallel
    {
        int nmax = OMAX(mx, OMAX(my, mz)) + 2 * S;
        double *b = (double *)malloc(sizeof(double) * nmax);
        /* x: BCxderVel */
#pragma omp for collapse(2) schedule(static)
        for (int k = 0; k < mz; k++)
            for (int j = 0; j < my; j++) {
                for (int q = 0; q < 3; q++) {
                    const double *f = fld[q];
                    for (int i = 0; i < mx; i++) b[i + S] = f[i + j * mx + k * mxy];
                    if (c->periodic_x) {
                        g_per(b, mx, S);
                    } else if (c->boundary_layer) { /* BCxderVel, src/boundary_condition_x.h:22-31 */
                        g_ext_top(b, mx, S);
                        g_odd_bot(b, S);
                        if (q == 0) g_perturb(o, b, S, j, k);
                    } else {
                        g_vel(b, mx, S);
                    }
                    for (int i = 0; i < mx; i++) o->g[q][i + j * mx + k * mxy] = dv1x(o, b, i + S, i);
                }
            }
        /* y: derDevV1yL, periodic */
#pragma omp for collapse(2) schedule(static)
        for (int k = 0; k < mz; k++)
            for (int i = 0; i < mx; i++) {
                for (int q = 0; q < 3; q++) {
                    const double *f = fld[q];
                    for (int j = 0; j < my; j++) b[j + SV] = f[i + j * mx + k * mxy];
                    g_per(b, my, SV);
                    for (int j = 0; j < my; j++)
                        o->g[3 + q][i + j * mx + k * mxy] = d1g(o->cVF, SV, o->d_dy, b, j + SV);
                }
            }
        /* z: derDevV1zL, periodic */
#pragma omp for collapse(2) schedule(static)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                for (int q = 0; q < 3; q++) {
                    const double *f = fld[q];
                    for (int k = 0; k < mz; k++) b[k + SV] = f[i + j * mx + k * mxy];
                    if (c->boundary_layer) { /* BCzderVel, src/boundary_condition_z.h:28-34 */
                        g_ext_top(b, mz, SV);
                        g_ext_bot(b, SV);
                    } else {
                        g_per(b, mz, SV);
                    }
                    for (int k = 0; k < mz; k++)
                        o->g[6 + q][i + j * mx + k * mxy] = d1g(o->cVF, SV, o->d_dz, b, k + SV);
                }
            }
        free(b);
    }
    long n = o->n;
#pragma omp parallel for schedule(static)
    for (long g = 0; g < n; g++) o->dil[g] = o->g[0][g] + o->g[4][g] + o->g[8][g];
}

/* ------------------------------------------------------------------ */
/* x part of the RHS (src/cuda_rhs.cu:9-128), writes out[] =            */
/* ------------------------------------------------------------------ */
static void rhs_x(orc *o, double **out) {
    const orc_config *c = &o->c;
    int mx = c->mx, my = c->my, mz = c->mz, S = o->S;
    long mxy = (long)mx * my;
    double Rgas = o->Rgas, gam = c->gam;
#pragma omp parallel
    {
        int L = mx + 2 * S;
        double *su = malloc(sizeof(double) * L), *sv_ = malloc(sizeof(double) * L), *sw = malloc(sizeof(double) * L);
        double *st = malloc(sizeof(double) * L), *sp = malloc(sizeof(double) * L), *sm = malloc(sizeof(double) * L);
        double *sl = malloc(sizeof(double) * L), *sd = malloc(sizeof(double) * L), *sr = malloc(sizeof(double) * L);
        double *sh = malloc(sizeof(double) * L);
#pragma omp for collapse(2) schedule(static)
        for (int k = 0; k < mz; k++)
            for (int j = 0; j < my; j++) {
                long base = j * mx + k * mxy;
                for (int i = 0; i < mx; i++) {
                    long g = base + i;
                    su[i + S] = o->u[g]; sv_[i + S] = o->v[g]; sw[i + S] = o->w[g];
                    st[i + S] = o->t[g]; sp[i + S] = o->p[g];
                    sm[i + S] = o->m[g]; sl[i + S] = o->l[g];
                    sd[i + S] = o->dil[g];
                    sr[i + S] = o->r[g]; sh[i + S] = o->h[g];
                }
                /* BCxNumber1 */
                if (c->periodic_x) {
                    g_per(su, mx, S); g_per(sv_, mx, S); g_per(sw, mx, S); g_per(st, mx, S);
                    g_per(sp, mx, S); g_per(sm, mx, S); g_per(sl, mx, S);
                } else if (c->boundary_layer) { /* src/boundary_condition_x.h:60-74 */
                    g_ext_top(su, mx, S); g_ext_top(sv_, mx, S); g_ext_top(sw, mx, S);
                    g_ext_top(sp, mx, S); g_ext_top(st, mx, S);
                    g_mir_bot(sp, S); g_mir_bot(st, S);
                    g_odd_bot(su, S); g_odd_bot(sv_, S); g_odd_bot(sw, S);
                    g_perturb(o, su, S, j, k);
                    for (int l = 0; l < S; l++) {
                        int gi[2] = {l, l + S + mx};
                        for (int q = 0; q < 2; q++) {
                            int ix = gi[q];
                            double suth = pow(st[ix], c->viscexp);
                            sm[ix] = suth / c->Re;
                            sl[ix] = suth / c->Re / c->Pr / o->Ec;
                        }
                    }
                } else {
                    g_mir(sp, mx, S);
                    g_vel(su, mx, S); g_vel(sv_, mx, S); g_vel(sw, mx, S);
                    g_ext(st, mx, S, c->twall_top, c->twall_bot);
                    /* mlBoundPT */
                    for (int l = 0; l < S; l++) {
                        int gi[2] = {l, l + S + mx};
                        for (int q = 0; q < 2; q++) {
                            int ix = gi[q];
                            double suth = pow(st[ix], c->viscexp);
                            sm[ix] = suth / c->Re;
                            sl[ix] = suth / c->Re / c->Pr / o->Ec;
                        }
                    }
                }
                /* BCxNumber2 */
                if (c->periodic_x) {
                    g_per(sd, mx, S);
                } else if (c->boundary_layer) {
                    g_ext_top(sd, mx, S);
                    g_mir_bot(sd, S);
                } else {
                    g_mir(sd, mx, S);
                }
                /* BCxNumber3 */
                if (c->periodic_x) {
                    g_per(sr, mx, S); g_per(sh, mx, S);
                } else {
                    for (int l = 0; l < S; l++) {
                        int gi[2] = {l, l + S + mx};
                        for (int q = 0; q < 2; q++) {
                            int ix = gi[q];
                            sh[ix] = st[ix] * Rgas * gam / (gam - 1.0) +
                                     0.5 * (su[ix] * su[ix] + sv_[ix] * sv_[ix] + sw[ix] * sw[ix]);
                            sr[ix] = sp[ix] / (Rgas * st[ix]);
                        }
                    }
                }
                for (int i = 0; i < mx; i++) {
                    long g = base + i;
                    int si = i + S;
                    double wrk1, wrk2;
                    double dudx = o->g[0][g], dvdx = o->g[1][g], dwdx = o->g[2][g];
                    double dudy = o->g[3][g], dudz = o->g[6][g];
                    double uX = (2 * dudx - 2. / 3. * o->dil[g]);
                    double vX = (dvdx + dudy);
                    double wX = (dwdx + dudz);
                    double eX = sm[si] * (uX * dudx + vX * dvdx + wX * dwdx);
                    wrk2 = dv1x(o, sm, si, i);
                    uX *= wrk2; vX *= wrk2; wX *= wrk2;
                    wrk1 = dv2x(o, su, si, i); uX = uX + wrk1 * sm[si];
                    wrk1 = dv2x(o, sv_, si, i); vX = vX + wrk1 * sm[si];
                    wrk1 = dv2x(o, sw, si, i); wX = wX + wrk1 * sm[si];
                    eX = eX + su[si] * uX + sv_[si] * vX + sw[si] * wX;
                    wrk1 = dv2x(o, st, si, i);
                    eX = eX + wrk1 * sl[si];
                    wrk2 = dv1x(o, sl, si, i);
                    wrk1 = dv1x(o, st, si, i);
                    eX = eX + wrk1 * wrk2;
                    wrk2 = dv1x(o, sd, si, i);
                    wrk1 = d1x(o, sp, si, i);
                    uX = uX + sm[si] * wrk2 / 3.0 - wrk1;
                    eX = eX + sm[si] * wrk2 / 3.0 * su[si];
                    double rX = fquad_x(o, sr, su, si, i);
                    uX = uX + fcube_x(o, sr, su, su, si, i);
                    vX = vX + fcube_x(o, sr, su, sv_, si, i);
                    wX = wX + fcube_x(o, sr, su, sw, si, i);
                    eX = eX + fcube_x(o, sr, su, sh, si, i);
                    out[0][g] = rX; out[1][g] = uX; out[2][g] = vX; out[3][g] = wX; out[4][g] = eX;
                }
            }
        free(su); free(sv_); free(sw); free(st); free(sp); free(sm); free(sl); free(sd); free(sr); free(sh);
    }
}

/* ------------------------------------------------------------------ */
/* y/z parts (src/cuda_rhs.cu:130-396), add +=                          */
/* dir 1: y (deviceRHSY), dir 2: z (deviceRHSZ incl. dpdz forcing)     */
/* ------------------------------------------------------------------ */
static void rhs_yz(orc *o, double **out, int dir) {
    const orc_config *c = &o->c;
    int mx = c->mx, my = c->my, mz = c->mz, S = o->S, SV = o->SV;
    long mxy = (long)mx * my;
    int n = (dir == 1) ? my : mz;
    long stride = (dir == 1) ? mx : mxy;
    int outer1 = (dir == 1) ? mz : my;
    double d = (dir == 1) ? o->d_dy : o->d_dz;
    double d2 = (dir == 1) ? o->d_d2y : o->d_d2z;
    double dpdz = o->dpdz;
#pragma omp parallel
    {
        int L = n + 2 * S;
        double *su = malloc(sizeof(double) * L), *sv_ = malloc(sizeof(double) * L), *sw = malloc(sizeof(double) * L);
        double *sm = malloc(sizeof(double) * L), *sd = malloc(sizeof(double) * L), *sp = malloc(sizeof(double) * L);
        double *sl = malloc(sizeof(double) * L), *st = malloc(sizeof(double) * L), *sr = malloc(sizeof(double) * L);
        double *sh = malloc(sizeof(double) * L);
#pragma omp for collapse(2) schedule(static)
        for (int a = 0; a < outer1; a++)
            for (int i = 0; i < mx; i++) {
                long base = (dir == 1) ? (i + (long)a * mxy) : (i + (long)a * mx);
                for (int q = 0; q < n; q++) {
                    long g = base + q * stride;
                    su[q + S] = o->u[g]; sv_[q + S] = o->v[g]; sw[q + S] = o->w[g];
                    sm[q + S] = o->m[g]; sd[q + S] = o->dil[g]; sp[q + S] = o->p[g];
                    sl[q + S] = o->l[g]; st[q + S] = o->t[g]; sr[q + S] = o->r[g]; sh[q + S] = o->h[g];
                }
                if (dir == 2 && c->boundary_layer) {
                    /* BCzNumber1-4 (src/boundary_condition_z.h): each staged
                       quantity extrapolated on its own about the end nodes */
                    double *all[10] = {su, sv_, sw, sm, sd, sp, sl, st, sr, sh};
                    for (int q = 0; q < 10; q++) {
                        g_ext_top(all[q], n, S);
                        g_ext_bot(all[q], S);
                    }
                } else {
                    g_per(su, n, S); g_per(sv_, n, S); g_per(sw, n, S); g_per(sm, n, S); g_per(sd, n, S);
                    g_per(sp, n, S); g_per(sl, n, S); g_per(st, n, S); g_per(sr, n, S); g_per(sh, n, S);
                }
                /* the convecting velocity and the diagonal component */
                const double *sa = (dir == 1) ? sv_ : sw;
                for (int q = 0; q < n; q++) {
                    long g = base + q * stride;
                    int si = q + S;
                    double wrk1, wrk2;
                    double rT, uT, vT, wT, eT;
                    if (dir == 1) {
                        double dvdx = o->g[1][g], dudy = o->g[3][g], dvdy = o->g[4][g], dwdy = o->g[5][g],
                               dvdz = o->g[7][g];
                        uT = (dudy + dvdx);
                        vT = (2 * dvdy - 2. / 3. * sd[si]);
                        wT = (dwdy + dvdz);
                        if (c->fix_y_work)
                            eT = sm[si] * (uT * dudy + vT * dvdy + wT * dwdy);
                        else
                            eT = sm[si] * (uT * dudy + vT * dvdy + wT * dvdz);
                    } else {
                        double dwdx = o->g[2][g], dwdy = o->g[5][g], dudz = o->g[6][g], dvdz = o->g[7][g],
                               dwdz = o->g[8][g];
                        uT = (dudz + dwdx);
                        vT = (dvdz + dwdy);
                        wT = (2 * dwdz - 2. / 3. * sd[si]);
                        eT = sm[si] * (uT * dudz + vT * dvdz + wT * dwdz);
                    }
                    wrk2 = d1g(o->cVF, SV, d, sm, si);
                    uT *= wrk2; vT *= wrk2; wT *= wrk2;
                    wrk1 = d2g(o->cVS, SV, d2, su, si); uT = uT + wrk1 * sm[si];
                    wrk1 = d2g(o->cVS, SV, d2, sv_, si); vT = vT + wrk1 * sm[si];
                    wrk1 = d2g(o->cVS, SV, d2, sw, si); wT = wT + wrk1 * sm[si];
                    eT = eT + su[si] * uT + sv_[si] * vT + sw[si] * wT;
                    /* dilatation derivative */
                    wrk2 = d1g(o->cVF, SV, d, sd, si);
                    if (dir == 1) {
                        vT = vT + sm[si] * wrk2 / 3.0;
                        eT = eT + sm[si] * wrk2 / 3.0 * sv_[si];
                    } else {
                        wT = wT + sm[si] * wrk2 / 3.0;
                        eT = eT + sm[si] * wrk2 / 3.0 * sw[si];
                    }
                    /* pressure */
                    wrk1 = d1g(o->cF, S, d, sp, si);
                    if (dir == 1) vT = vT - wrk1; else wT = wT - wrk1;
                    /* Fourier */
                    wrk1 = d2g(o->cVS, SV, d2, st, si);
                    eT = eT + wrk1 * sl[si];
                    wrk2 = d1g(o->cVF, SV, d, sl, si);
                    wrk1 = d1g(o->cVF, SV, d, st, si);
                    eT = eT + wrk1 * wrk2;
                    /* split-form convective fluxes */
                    rT = fquad_g(o->cF, S, d, sr, sa, si);
                    uT = uT + fcube_g(o->cF, S, d, sr, sa, su, si);
                    vT = vT + fcube_g(o->cF, S, d, sr, sa, sv_, si);
                    wT = wT + fcube_g(o->cF, S, d, sr, sa, sw, si);
                    eT = eT + fcube_g(o->cF, S, d, sr, sa, sh, si);
                    out[0][g] += rT;
                    out[1][g] += uT;
                    out[2][g] += vT;
                    if (dir == 2) {
                        out[3][g] += wT + dpdz;
                        out[4][g] += eT + dpdz * sw[si];
                    } else {
                        out[3][g] += wT;
                        out[4][g] += eT;
                    }
                }
            }
        free(su); free(sv_); free(sw); free(sm); free(sd); free(sp); free(sl); free(st); free(sr); free(sh);
    }
}

/* src/sponge.cu:31-41 with references held as conserved variables */
static void add_sponge(orc *o, double **out) {
    const orc_config *c = &o->c;
    int mx = c->mx, my = c->my, mz = c->mz;
    long mxy = (long)mx * my;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++)
            for (int i = 0; i < mx; i++) {
                long g = i + j * mx + k * mxy;
                long i2 = i + (long)k * mx;
                double sg = (o->spx[i] + o->spz[k]);
                out[0][g] += sg * (o->ref[0][i2] - o->r[g]);
                out[1][g] += sg * (o->ref[1][i2] - o->r[g] * o->u[g]);
                out[2][g] += sg * (o->ref[2][i2] - o->r[g] * o->v[g]);
                out[3][g] += sg * (o->ref[3][i2] - o->r[g] * o->w[g]);
                out[4][g] += sg * (o->ref[4][i2] - o->e[g]);
            }
}

/* src/cuda_main.cu:15-42 */
static void calc_rhs(orc *o, double **out) {
    calc_state(o);
    der_vel(o);
    rhs_x(o, out);
    rhs_yz(o, out, 1);
    rhs_yz(o, out, 2);
    if (o->c.sponge) add_sponge(o, out);
}

/* ------------------------------------------------------------------ */
/* reductions (src/cuda_math.cu:143-230)                               */
/* ------------------------------------------------------------------ */
static double vol_integral(const orc *o, const double *f, int average) {
    const orc_config *c = &o->c;
    int mx = c->mx;
    long n = o->n;
    double sum = 0.0;
#pragma omp parallel for reduction(+ : sum) schedule(static)
    for (long it = 0; it < n; it++) {
        int i = (int)(it % mx);
        if (average)
            sum += f[it] * o->dxv[i] / o->d_dy / o->d_dz / c->Lx / c->Ly / c->Lz;
        else
            sum += f[it] * o->dxv[i] / o->d_dy / o->d_dz;
    }
    return sum;
}

/* src/calc_stress.cu:140-160 + hostReduceToMin; mu is the field left by
 * the most recent calcState (src/cuda_main.cu:252 reads d_m). */
static double calc_dt(const orc *o) {
    const orc_config *c = &o->c;
    int mx = c->mx;
    long n = o->n;
    double gam = c->gam;
    double best = 100000;
#pragma omp parallel for reduction(min : best) schedule(static)
    for (long g = 0; g < n; g++) {
        int i = (int)(g % mx);
        double ien = o->e[g] / o->r[g] - 0.5 * (o->u[g] * o->u[g] + o->v[g] * o->v[g] + o->w[g] * o->w[g]);
        double sos = pow(gam * (gam - 1) * ien, 0.5);
        double dx = o->dxv[i];
        double d2x = dx * dx;
        double cv1 = (fabs(o->u[g]) + sos) / dx;
        double cv2 = (fabs(o->v[g]) + sos) * o->d_dy;
        double cv3 = (fabs(o->w[g]) + sos) * o->d_dz;
        double conv = OMAX(cv1, OMAX(cv2, cv3));
        double vs1 = o->m[g] / d2x, vs2 = o->m[g] * o->d_d2y, vs3 = o->m[g] * o->d_d2z;
        double visc = OMAX(vs1, OMAX(vs2, vs3));
        double val = c->cfl / OMAX(conv, visc);
        best = OMIN(best, val);
    }
    return best;
}

/* src/calc_stress.cu:98-120 */
static void calc_pressure_grad(orc *o) {
    long n = o->n;
    double *wk = fal(n);
#pragma omp parallel for schedule(static)
    for (long g = 0; g < n; g++) wk[g] = o->r[g] * o->w[g];
    double prev = o->dpdz;
    double rbulk = vol_integral(o, o->r, 0);
    double a = vol_integral(o, wk, 0);
    o->dpdz = 0.99 * (prev) - 0.5 * (a / (rbulk) - 1);
    free(wk);
}

/* src/calc_stress.cu:162-201 */
static void calc_bulk(orc *o, double *par1, double *par2) {
    long n = o->n;
    double *wk = fal(n);
    if (o->c.forcing) {
        double rbulk = vol_integral(o, o->r, 0);
#pragma omp parallel for schedule(static)
        for (long g = 0; g < n; g++) wk[g] = o->r[g] * o->w[g];
        *par1 = vol_integral(o, wk, 0) / rbulk;
        *par2 = vol_integral(o, o->e, 0);
    } else {
#pragma omp parallel for schedule(static)
        for (long g = 0; g < n; g++) wk[g] = o->u[g] * o->u[g] + o->v[g] * o->v[g] + o->w[g] * o->w[g];
        *par1 = vol_integral(o, wk, 1);
        *par2 = NAN;
    }
    free(wk);
}

/* ------------------------------------------------------------------ */
/* RK stage updates, shared by the solver loop and the scalar ODE test */
/* ------------------------------------------------------------------ */
static const double smr_alpha[3] = {0., -17. / 60., -5. / 12.};
static const double smr_beta[3] = {8. / 15., 5. / 12., 3. / 4.};

/* sumLowStorageRK3 (src/cuda_main.cu:244-247) */
static void upd_smr(double *var, const double *a, const double *b, double dt, int st, long n) {
#pragma omp parallel for schedule(static)
    for (long g = 0; g < n; g++) var[g] = var[g] + (dt) * (smr_alpha[st] * a[g] + smr_beta[st] * b[g]);
}
/* eulerSum / eulerSum3 / rk3final (src/cuda_main.cu:188-216); dv==NULL
 * means no division by density */
static void upd_kutta(double *a, const double *b, const double *k1, const double *k2, const double *k3,
                      const double *dv, double dt, int st, long n) {
#pragma omp parallel for schedule(static)
    for (long g = 0; g < n; g++) {
        double val;
        if (st == 0) val = (b[g] + k1[g] * (dt) / 2.0);
        else if (st == 1) val = b[g] + (2 * k2[g] - k1[g]) * (dt);
        else val = b[g] + (dt) * (k1[g] + 4 * k2[g] + k3[g]) / 6.;
        a[g] = dv ? val / dv[g] : val;
    }
}
/* classic RK4 (no reference counterpart) */
static void upd_rk4(double *a, const double *q0, const double *k1, const double *k2, const double *k3,
                    const double *k4, const double *dv, double dt, int st, long n) {
#pragma omp parallel for schedule(static)
    for (long g = 0; g < n; g++) {
        double val;
        if (st == 0) val = q0[g] + (0.5 * dt) * k1[g];
        else if (st == 1) val = q0[g] + (0.5 * dt) * k2[g];
        else if (st == 2) val = q0[g] + dt * k3[g];
        else val = q0[g] + (dt / 6.0) * (k1[g] + 2.0 * k2[g] + 2.0 * k3[g] + k4[g]);
        a[g] = dv ? val / dv[g] : val;
    }
}

static void vmul(double *a, const double *b, const double *c, long n) {
#pragma omp parallel for schedule(static)
    for (long g = 0; g < n; g++) a[g] = b[g] * c[g];
}
static void vdiv(double *a, const double *b, const double *c, long n) {
#pragma omp parallel for schedule(static)
    for (long g = 0; g < n; g++) a[g] = b[g] / c[g];
}

/* ------------------------------------------------------------------ */
/* public API                                                          */
/* ------------------------------------------------------------------ */
void orc_set_state(orc *o, const double *r, const double *u, const double *v, const double *w, const double *e) {
    size_t b = sizeof(double) * (size_t)o->n;
    memcpy(o->r, r, b); memcpy(o->u, u, b); memcpy(o->v, v, b); memcpy(o->w, w, b); memcpy(o->e, e, b);
    calc_state(o); /* copyField(0) ends with calcState (src/cuda_utils.cu:329) */
}

void orc_get_state(const orc *o, double *r, double *u, double *v, double *w, double *e) {
    size_t b = sizeof(double) * (size_t)o->n;
    memcpy(r, o->r, b); memcpy(u, o->u, b); memcpy(v, o->v, b); memcpy(w, o->w, b); memcpy(e, o->e, b);
}

/* one RHS evaluation of the current state into out[5] (host arrays) */
void orc_rhs(orc *o, double *rr, double *ru, double *rv, double *rw, double *re) {
    double *out[5] = {rr, ru, rv, rw, re};
    calc_rhs(o, out);
}

/* internal fields for tests: 0..4 h,t,p,m,l; 5..13 gradients; 14 dil */
void orc_get_field(const orc *o, int id, double *out) {
    const double *src = NULL;
    switch (id) {
    case 0: src = o->h; break;
    case 1: src = o->t; break;
    case 2: src = o->p; break;
    case 3: src = o->m; break;
    case 4: src = o->l; break;
    case 14: src = o->dil; break;
    default:
        if (id >= 5 && id <= 13) src = o->g[id - 5];
    }
    if (src) memcpy(out, src, sizeof(double) * (size_t)o->n);
}

/* grid: x, xp, xpp, y, z, dxv and scalars
 * sc = {dx, d_dx, d_dy, d_dz, d_d2x, d_d2y, d_d2z, Rgas, Ec} */
void orc_get_grid(const orc *o, double *x, double *xp, double *xpp, double *y, double *z, double *dxv, double *sc,
                  double *vsx) {
    int mx = o->c.mx;
    memcpy(x, o->x, sizeof(double) * mx); memcpy(xp, o->xp, sizeof(double) * mx);
    memcpy(xpp, o->xpp, sizeof(double) * mx); memcpy(dxv, o->dxv, sizeof(double) * mx);
    memcpy(y, o->y, sizeof(double) * o->c.my); memcpy(z, o->z, sizeof(double) * o->c.mz);
    sc[0] = o->dx; sc[1] = o->d_dx; sc[2] = o->d_dy; sc[3] = o->d_dz;
    sc[4] = o->d_d2x; sc[5] = o->d_d2y; sc[6] = o->d_d2z; sc[7] = o->Rgas; sc[8] = o->Ec;
    if (vsx) memcpy(vsx, o->VSx, sizeof(double) * mx * (2 * o->SV + 1));
}

void orc_set_sponge(orc *o, const double *sx, const double *sz, const double *rr, const double *ru, const double *rv,
                    const double *rw, const double *re) {
    memcpy(o->spx, sx, sizeof(double) * o->c.mx);
    memcpy(o->spz, sz, sizeof(double) * o->c.mz);
    const double *src[5] = {rr, ru, rv, rw, re};
    for (int q = 0; q < 5; q++) memcpy(o->ref[q], src[q], sizeof(double) * o->c.mx * o->c.mz);
}

double orc_get_dt(const orc *o) { return o->dt; }
double orc_get_dpdz(const orc *o) { return o->dpdz; }
double orc_get_time(const orc *o) { return o->time; }
void orc_set_time(orc *o, double t) { o->time = t; }
void orc_set_dt(orc *o, double dt) { o->dt = dt; }
double orc_calc_dt(const orc *o) { return calc_dt(o); }
void orc_bulk(orc *o, double *par1, double *par2) { calc_bulk(o, par1, par2); }
double orc_vol_integral(const orc *o, const double *f, int average) { return vol_integral(o, f, average); }

/* runSimulation[LowStorage] (src/cuda_main.cu:44-186); par1/par2 are
 * written at bulk-check steps and NaN elsewhere; time[istep] is the
 * simulation time after the step. */
void orc_step(orc *o, int nsteps, double *par1, double *par2, double *time) {
    const orc_config *c = &o->c;
    long n = o->n;
    double *st[5] = {o->r, o->u, o->v, o->w, o->e};
    for (int istep = 0; istep < nsteps; istep++) {
        if (istep % c->check_cfl == 0) {
            o->dt = calc_dt(o);
            if (c->forcing) calc_pressure_grad(o);
        }
        double dt = o->dt;
        o->time += dt;
        if (time) time[istep] = o->time;
        if (par1) par1[istep] = NAN;
        if (par2) par2[istep] = NAN;
        if (istep % c->check_bulk == 0) {
            double a, b;
            calc_bulk(o, &a, &b);
            if (par1) par1[istep] = a;
            if (par2) par2[istep] = b;
        }
        if (c->scheme == 0) {
            for (int s = 0; s < 3; s++) {
                double **rnew = (s == 1) ? o->R2 : o->R1;
                double **rold = (s == 2) ? o->R2 : o->R1;
                calc_rhs(o, rnew);
                /* deviceMul: momentum with the density of this stage's input */
                vmul(o->u, o->u, o->r, n); vmul(o->v, o->v, o->r, n); vmul(o->w, o->w, o->r, n);
                for (int q = 0; q < 5; q++) upd_smr(st[q], rold[q], rnew[q], dt, s, n);
                vdiv(o->u, o->u, o->r, n); vdiv(o->v, o->v, o->r, n); vdiv(o->w, o->w, o->r, n);
            }
        } else if (c->scheme == 1) {
            vmul(o->O[1], o->r, o->u, n); vmul(o->O[2], o->r, o->v, n); vmul(o->O[3], o->r, o->w, n);
            memcpy(o->O[0], o->r, sizeof(double) * n);
            memcpy(o->O[4], o->e, sizeof(double) * n);
            double **K[3] = {o->R1, o->R2, o->R3};
            for (int s = 0; s < 3; s++) {
                calc_rhs(o, K[s]);
                upd_kutta(o->r, o->O[0], o->R1[0], o->R2[0], o->R3[0], NULL, dt, s, n);
                upd_kutta(o->e, o->O[4], o->R1[4], o->R2[4], o->R3[4], NULL, dt, s, n);
                for (int q = 1; q < 4; q++) upd_kutta(st[q], o->O[q], o->R1[q], o->R2[q], o->R3[q], o->r, dt, s, n);
            }
        } else {
            /* classic RK4 on conserved variables: k1..k4 in R1,R2,R3,O-side buffer */
            double *q0[5];
            for (int q = 0; q < 5; q++) q0[q] = fal(n);
            memcpy(q0[0], o->r, sizeof(double) * n);
            vmul(q0[1], o->r, o->u, n); vmul(q0[2], o->r, o->v, n); vmul(q0[3], o->r, o->w, n);
            memcpy(q0[4], o->e, sizeof(double) * n);
            double **K[4] = {o->R1, o->R2, o->R3, o->O};
            for (int s = 0; s < 4; s++) {
                calc_rhs(o, K[s]);
                upd_rk4(o->r, q0[0], o->R1[0], o->R2[0], o->R3[0], o->O[0], NULL, dt, s, n);
                upd_rk4(o->e, q0[4], o->R1[4], o->R2[4], o->R3[4], o->O[4], NULL, dt, s, n);
                for (int q = 1; q < 4; q++)
                    upd_rk4(st[q], q0[q], o->R1[q], o->R2[q], o->R3[q], o->O[q], o->r, dt, s, n);
            }
            for (int q = 0; q < 5; q++) free(q0[q]);
        }
    }
}

/* Scalar ODE y' = lam*y through the same stage-update routines as the
 * solver loop: one step of the given scheme, returns y1 for y0 = 1. */
double orc_rk_scalar(int scheme, double lam, double dt) {
    double y = 1.0, q0 = 1.0;
    double k[4] = {0, 0, 0, 0};
    if (scheme == 0) {
        double r1 = 0, r2 = 0;
        for (int s = 0; s < 3; s++) {
            double *rnew = (s == 1) ? &r2 : &r1;
            double *rold = (s == 2) ? &r2 : &r1;
            *rnew = lam * y;
            upd_smr(&y, rold, rnew, dt, s, 1);
        }
    } else if (scheme == 1) {
        for (int s = 0; s < 3; s++) {
            k[s] = lam * y;
            upd_kutta(&y, &q0, &k[0], &k[1], &k[2], NULL, dt, s, 1);
        }
    } else {
        for (int s = 0; s < 4; s++) {
            k[s] = lam * y;
            upd_rk4(&y, &q0, &k[0], &k[1], &k[2], &k[3], NULL, dt, s, 1);
        }
    }
    return y;
}

/* periodic pencil primitives (the y/z forms used by the RHS) for tests */
static void per_buf(const double *f, double *b, int n, int S) {
    for (int i = 0; i < n; i++) b[i + S] = f[i];
    g_per(b, n, S);
}
void orc_pencil_d1(int n, int s, double d, const double *f, double *out) {
    double cF[4], cS[5];
    orc_coeffs(s, cF, cS);
    double *b = malloc(sizeof(double) * (n + 2 * s));
    per_buf(f, b, n, s);
    for (int i = 0; i < n; i++) out[i] = d1g(cF, s, d, b, i + s);
    free(b);
}
void orc_pencil_d2(int n, int s, double d2, const double *f, double *out) {
    double cF[4], cS[5];
    orc_coeffs(s, cF, cS);
    double *b = malloc(sizeof(double) * (n + 2 * s));
    per_buf(f, b, n, s);
    for (int i = 0; i < n; i++) out[i] = d2g(cS, s, d2, b, i + s);
    free(b);
}
void orc_pencil_flux(int n, int s, double d, const double *f, const double *g, const double *h, double *out) {
    double cF[4], cS[5];
    orc_coeffs(s, cF, cS);
    double *bf = malloc(sizeof(double) * (n + 2 * s)), *bg = malloc(sizeof(double) * (n + 2 * s));
    double *bh = malloc(sizeof(double) * (n + 2 * s));
    per_buf(f, bf, n, s);
    per_buf(g, bg, n, s);
    if (h) per_buf(h, bh, n, s);
    for (int i = 0; i < n; i++)
        out[i] = h ? fcube_g(cF, s, d, bf, bg, bh, i + s) : fquad_g(cF, s, d, bf, bg, i + s);
    free(bf); free(bg); free(bh);
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
