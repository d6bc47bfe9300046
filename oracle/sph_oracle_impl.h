/*
 * sph_oracle_impl.h -- body of the CPU oracle (TEST INFRASTRUCTURE ONLY), compiled twice by
 * sph_oracle.c: REAL=float with S=f32, REAL=double with S=f64. Not a public header.
 *
 * The physics follows, per function:
 *   kernels          include/spearhed/sph/CubicSplineKernel.hpp:47-132, QuinticSplineKernel.hpp:46-160
 *                    (W, dWdr, gradWScalar; the guard !(r > 0) || r >= support*h -> 0), plus
 *                    Wendland C2 (Wendland 1995; Dehnen & Aly 2012 MNRAS 425:1068 eq. 5 with H = 2h).
 *   pair geometry    packages/pmacc/include/spmacc/particles/algorithms/ParticleParticleInteraction.hpp:236-311
 *                    (pre-shifted staged position, strict r2 < R^2 cull, invR = rsqrt(r2) only
 *                    when r2 > 0, self pair excluded by identity).
 *   density          include/spearhed/particles/density/DensitySummation.hpp:50-94 (gather with own h,
 *                    self term m_i W(0, h_i), coincident non-self pairs contribute m_j W(0, h_i)).
 *   hydro force      include/spearhed/sph/HydroForces.hpp:112-175 (symmetric pressure gradient,
 *                    P dV energy with the own-side P_i/rho_i^2), EquationOfState.hpp:32-47.
 *   viscosity        Monaghan 1997 JCP 136:298 / Monaghan & Gingold 1983 (alpha, beta; mu with
 *                    eta^2 = 0.01 h^2; arithmetic means of h, c, rho).
 *   SPMHD            Price 2012 JCP 231:759 sec. 4 (Maxwell-stress momentum equation, B induction
 *                    with the own-side kernel gradient, difference divergence of B), Borve et al.
 *                    2001 tensile correction, Tricco & Price 2012 JCP 231:7214 (constrained
 *                    hyperbolic cleaning: symmetric grad psi, psi' = -c_h^2 div B - psi/tau
 *                    - psi div v / 2).
 *   drift / kick     include/spearhed/particles/pusher/PushVelocity.hpp:35-42,
 *                    EulerIntegrate.hpp:43-52 (symplectic Euler: drift with the old v, kick after
 *                    the force pass), extended to B and psi.
 *   step order       include/spearhed/control/Simulation.x.cpp:196-310.
 */

#define ORC_CAT2(a, b) a##_##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define FN(name) ORC_CAT(name, S)

#define RA(p) ((REAL*)(p))

static const REAL FN(kPi) = (REAL)3.14159265358979323846;

/* Kernel normalisation sigma in 3D (CubicSplineKernel.hpp:47-58, QuinticSplineKernel.hpp:46-57). */
static inline REAL FN(knorm)(int kernel)
{
    if (kernel == SPH_KERNEL_CUBIC) return (REAL)1 / FN(kPi);
    if (kernel == SPH_KERNEL_QUINTIC) return (REAL)1 / ((REAL)120 * FN(kPi));
    return (REAL)21 / ((REAL)16 * FN(kPi));
}

/* Shape function f(q), W = sigma/h^3 f(q) (CubicSplineKernel.hpp:67-86,
 * QuinticSplineKernel.hpp:66-101). */
static inline REAL FN(kshape)(int kernel, REAL q)
{
    if (kernel == SPH_KERNEL_CUBIC) {
        if (q < (REAL)1) return (REAL)1 - (REAL)1.5 * q * q + (REAL)0.75 * q * q * q;
        if (q < (REAL)2) {
            REAL t = (REAL)2 - q;
            return (REAL)0.25 * t * t * t;
        }
        return (REAL)0;
    }
    if (kernel == SPH_KERNEL_QUINTIC) {
        REAL a = (REAL)3 - q, b = (REAL)2 - q, c = (REAL)1 - q;
        if (q < (REAL)1) return a * a * a * a * a - (REAL)6 * b * b * b * b * b + (REAL)15 * c * c * c * c * c;
        if (q < (REAL)2) return a * a * a * a * a - (REAL)6 * b * b * b * b * b;
        if (q < (REAL)3) return a * a * a * a * a;
        return (REAL)0;
    }
    /* Wendland C2: (1 - q/2)^4 (1 + 2q) for q < 2 */
    if (q < (REAL)2) {
        REAL t = (REAL)1 - (REAL)0.5 * q;
        REAL t2 = t * t;
        return t2 * t2 * ((REAL)1 + (REAL)2 * q);
    }
    return (REAL)0;
}

/* Radial-derivative shape g(q), dW/dr = sigma/h^4 g(q) (CubicSplineKernel.hpp:94-114,
 * QuinticSplineKernel.hpp:109-143). Wendland: g = -5 q (1 - q/2)^3. */
static inline REAL FN(kdshape)(int kernel, REAL q)
{
    if (kernel == SPH_KERNEL_CUBIC) {
        if (q >= (REAL)0 && q < (REAL)1) return (REAL)2.25 * q * q - (REAL)3 * q;
        if (q < (REAL)2) {
            REAL t = (REAL)2 - q;
            return -(REAL)0.75 * t * t;
        }
        return (REAL)0;
    }
    if (kernel == SPH_KERNEL_QUINTIC) {
        REAL a = (REAL)3 - q, b = (REAL)2 - q, c = (REAL)1 - q;
        if (q < (REAL)1) return -(REAL)5 * a * a * a * a + (REAL)30 * b * b * b * b - (REAL)75 * c * c * c * c;
        if (q < (REAL)2) return -(REAL)5 * a * a * a * a + (REAL)30 * b * b * b * b;
        if (q < (REAL)3) return -(REAL)5 * a * a * a * a;
        return (REAL)0;
    }
    if (q < (REAL)2) {
        REAL t = (REAL)1 - (REAL)0.5 * q;
        return -(REAL)5 * q * t * t * t;
    }
    return (REAL)0;
}

/* Per-particle kernel factors: W = wfac f(q); gradient scalar s (grad_i W = s r_ij) is
 * gfac g(q) invR for the splines and gfac (1 - q/2)^3 for Wendland C2 (gfac = -5 sigma/h^5),
 * which avoids the 1/r of gradWScalar (CubicSplineKernel.hpp:129-137) exactly. */
static inline void FN(kfactors)(int kernel, REAL h, REAL* invh, REAL* wfac, REAL* gfac)
{
    REAL sig = FN(knorm)(kernel);
    *invh = (REAL)1 / h;
    *wfac = sig / (h * h * h);
    if (kernel == SPH_KERNEL_WENDLAND)
        *gfac = -(REAL)5 * sig / (h * h * h * h * h);
    else
        *gfac = sig / (h * h * h * h);
}

/* gradWScalar with the reference guard (!(r > 0) || r >= support h -> 0), in q = r * (1/h) form (the
 * reference divides x = r / h per pair; the results agree within FP32 rounding). */
static inline REAL FN(kgrad)(int kernel, REAL rr, REAL invR, REAL q, REAL gfac)
{
    REAL sup = (REAL)ksupport(kernel);
    if (!(rr > (REAL)0) || !(q < sup)) return (REAL)0;
    if (kernel == SPH_KERNEL_WENDLAND) {
        REAL t = (REAL)1 - (REAL)0.5 * q;
        return gfac * (t * t * t);
    }
    return gfac * FN(kdshape)(kernel, q) * invR;
}

double FN(orc_kernel_w)(int32_t kernel, double r, double h)
{
    REAL invh, wfac, gfac;
    FN(kfactors)(kernel, (REAL)h, &invh, &wfac, &gfac);
    REAL q = (REAL)r * invh;
    return (double)(wfac * FN(kshape)(kernel, q));
}

double FN(orc_kernel_dwdr)(int32_t kernel, double r, double h)
{
    REAL invh, wfac, gfac;
    FN(kfactors)(kernel, (REAL)h, &invh, &wfac, &gfac);
    REAL q = (REAL)r * invh;
    REAL sig = FN(knorm)(kernel);
    REAL g4 = sig / ((REAL)h * (REAL)h * (REAL)h * (REAL)h);
    return (double)(g4 * FN(kdshape)(kernel, q));
}

/* Pair coordinates used by the interaction passes: FP32 absolute (plain), FP32 relative to the
 * particle's own supercell origin (mixed), or FP64 absolute (f64). Mirrors relativePos + region
 * origin (packages/pmacc/include/spmacc/particles/attributes/RelativePosition.hpp:36-37). */
static void FN(pair_coords)(const orc_grid* g, int32_t prec, const orc_state* st, REAL* pc[3])
{
    int64_t n = st->n;
    for (int a = 0; a < 3; ++a) {
        for (int64_t i = 0; i < n; ++i) {
            double x = st->pos[a][i];
            if (prec == ORC_PREC_MIXED) {
                double t = (x - g->lo[a]) * g->inv_s[a];
                int c = (int)fmin(fmax(floor(t), 0.0), (double)(g->n[a] - 1));
                pc[a][i] = (REAL)((x - g->lo[a]) - (double)c * g->s[a]);
            } else {
                pc[a][i] = (REAL)x;
            }
        }
    }
}

/* Neighbour supercell (cell + d) with its staged-position shift; returns 0 if outside a
 * non-periodic axis. */
static inline int FN(nb_cell)(const orc_grid* g, int32_t prec, const sph_run_config* cfg,
                              const int c[3], const int d[3], int* cc, REAL shift[3])
{
    int idx[3];
    for (int a = 0; a < 3; ++a) {
        int x = c[a] + d[a];
        int wrap = 0;
        if (cfg->periodic[a]) {
            /* floor division: a stencil reaching 2 cells may wrap a 1-cell axis twice */
            wrap = x >= 0 ? x / g->n[a] : -((-x + g->n[a] - 1) / g->n[a]);
            x -= wrap * g->n[a];
        } else if (x < 0 || x >= g->n[a]) {
            return 0;
        }
        idx[a] = x;
        if (prec == ORC_PREC_MIXED)
            shift[a] = (REAL)d[a] * (REAL)g->s32[a];
        else if (prec == ORC_PREC_PLAIN)
            shift[a] = (REAL)wrap * (REAL)g->L32[a];
        else
            shift[a] = (REAL)((double)wrap * g->L[a]);
    }
    *cc = idx[0] + g->n[0] * (idx[1] + g->n[1] * idx[2]);
    return 1;
}

/* Squared distance of own pair coordinate to a staged neighbour: dx^2 + dy^2 + dz^2
 * (ParticleParticleInteraction.hpp:281-289), evaluated as fma(dz, dz, fma(dy, dy, dx*dx)), the
 * GPU kernels' evaluation (the reference leaves the contraction to its compiler). */
static inline REAL FN(geom)(const REAL pi[3], REAL* const pc[3], int64_t j, const REAL shift[3], REAL d[3])
{
    for (int a = 0; a < 3; ++a) {
        REAL staged = pc[a][j] + shift[a];
        d[a] = pi[a] - staged;
    }
    return RFMA(d[2], d[2], RFMA(d[1], d[1], d[0] * d[0]));
}

/* A run of source slots [b, e) staged with one shift: a stencil supercell (binned traversal) or a
 * neighbour region under one periodic image (region traversal). `own` marks the target's own
 * supercell / region with no shift, where the target itself is skipped (the reference's isSelf,
 * ParticleParticleInteraction.hpp:299). */
typedef struct FN(span) {
    int64_t b, e;
    int own;
    REAL sh[3];
} FN(span);

/* The stencil of one cell as spans, (2 reach + 1)^3 supercells x fastest (fewer at a non-periodic
 * wall): 27 for supercells of edge >= H, 125 for edge >= H/2. */
static int FN(cell_spans)(const orc_grid* g, int32_t prec, const sph_run_config* cfg, const uint32_t* cell_start,
                          int cell, FN(span)* sp)
{
    int c[3];
    cell_coords(g, cell, c);
    int n = 0;
    const int r = g->reach;
    for (int dz = -r; dz <= r; ++dz)
        for (int dy = -r; dy <= r; ++dy)
            for (int dx = -r; dx <= r; ++dx) {
                int d[3] = {dx, dy, dz};
                int cc;
                if (FN(nb_cell)(g, prec, cfg, c, d, &cc, sp[n].sh)) {
                    sp[n].b = cell_start[cc];
                    sp[n].e = cell_start[cc + 1];
                    sp[n].own = (dx == 0 && dy == 0 && dz == 0);
                    ++n;
                }
            }
    return n;
}

int FN(orc_permute)(orc_state* st, const int64_t* perm)
{
    int64_t n = st->n;
    void* tmp = malloc((size_t)n * 8 + 16);
    if (!tmp) return -2;
#define PERMUTE(ptr, T)                                                  \
    if (ptr) {                                                           \
        T* src = (T*)(ptr);                                              \
        T* t = (T*)tmp;                                                  \
        for (int64_t i = 0; i < n; ++i) t[i] = src[perm[i]];             \
        memcpy(src, t, (size_t)n * sizeof(T));                           \
    }
    PERMUTE(st->id, uint64_t)
    PERMUTE(st->mask, uint8_t)
    PERMUTE(st->species, uint8_t)
    for (int a = 0; a < 3; ++a) {
        PERMUTE(st->pos[a], double)
        PERMUTE(st->vel[a], REAL)
        PERMUTE(st->B[a], REAL)
        PERMUTE(st->dvdt[a], REAL)
        PERMUTE(st->dbdt[a], REAL)
    }
    PERMUTE(st->mass, REAL)
    PERMUTE(st->rho, REAL)
    PERMUTE(st->h, REAL)
    PERMUTE(st->u, REAL)
    PERMUTE(st->psi, REAL)
    PERMUTE(st->dudt, REAL)
    PERMUTE(st->dpsidt, REAL)
    PERMUTE(st->divb, REAL)
    PERMUTE(st->vsig, REAL)
    PERMUTE(st->nnb, uint32_t)
#undef PERMUTE
    free(tmp);
    return 0;
}

/* Supercell re-sort (replaces UpdateVolumes / the region AABB refit,
 * packages/pmacc/include/spmacc/particles/regions/RegionBoundsUpdate.hpp:39-214). */
int FN(orc_resort)(const sph_run_config* cfg, int32_t prec, orc_state* st, uint32_t* cell_start)
{
    orc_grid g;
    if (orc_make_grid(cfg, &g)) return -1;
    int64_t n = st->n;
    uint32_t* key = (uint32_t*)malloc((size_t)n * 4 + 4);
    int64_t* perm = (int64_t*)malloc((size_t)n * 8 + 8);
    if (!key || !perm) return -2;
    orc_bin(cfg, prec, n, st->pos, key);
    orc_sort(n, key, st->id, g.ncell, perm, cell_start);
    FN(orc_permute)(st, perm);
    free(key);
    free(perm);
    return 0;
}

typedef struct FN(prim) {
    REAL invh, wfac, gfac, c, pr, ph, bdr, psr;
} FN(prim);

/* Per-particle derived quantities after the density pass: EOS (EquationOfState.hpp:32-47),
 * P/rho^2 (HydroForces.hpp:112-121, 124-137), fast magnetosonic bound c = sqrt((gamma P + B^2)/rho). */
static void FN(prims)(const sph_run_config* cfg, const orc_state* st, FN(prim)* pr)
{
    int64_t n = st->n;
    REAL gam = (REAL)cfg->gamma;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        FN(prim) p;
        memset(&p, 0, sizeof(p));
        if (is_source(st->species[i], st->mask[i])) {
            REAL h = RA(st->h)[i];
            FN(kfactors)(cfg->kernel, h, &p.invh, &p.wfac, &p.gfac);
            REAL rho = RA(st->rho)[i];
            REAL u = RA(st->u)[i];
            REAL P = (gam - (REAL)1) * rho * u;
            REAL invrho = (REAL)1 / rho;
            REAL B2 = (REAL)0;
            if (cfg->mhd) {
                REAL bx = RA(st->B[0])[i], by = RA(st->B[1])[i], bz = RA(st->B[2])[i];
                B2 = bx * bx + by * by + bz * bz;
            }
            p.bdr = invrho * invrho;
            p.ph = P * p.bdr;
            p.pr = (P + (REAL)0.5 * B2) * p.bdr;
            p.psr = cfg->mhd ? RA(st->psi)[i] * p.bdr : (REAL)0;
            p.c = (REAL)sqrt((gam * P + B2) * invrho);
        }
        pr[i] = p;
    }
}

/* Density of target i summed over the source spans (DensitySummation.hpp:50-138 semantics): the
 * self term plus m_j W(r_ij, h_i) over sources within H, with the neighbour count. sigma/h^3 is
 * factored out of the sum and the self term added after the neighbours; the reference seeds rho with
 * the self term and multiplies each pair by sigma/h^3 (same value within FP32 rounding). */
static void FN(density_one)(const sph_run_config* cfg, orc_state* st, REAL* const pc[3], REAL H2, REAL f0, int64_t i,
                            const FN(span)* sp, int nsp)
{
    REAL* rho = RA(st->rho);
    const REAL* m = RA(st->mass);
    REAL invh, wfac, gfac;
    FN(kfactors)(cfg->kernel, RA(st->h)[i], &invh, &wfac, &gfac);
    REAL pi[3] = {pc[0][i], pc[1][i], pc[2][i]};
    REAL acc = (REAL)0;
    double sc = 0.0;
    uint32_t cnt = 0;
    for (int k = 0; k < nsp; ++k) {
        const int own = sp[k].own;
        for (int64_t j = sp[k].b; j < sp[k].e; ++j) {
            if (!is_source(st->species[j], st->mask[j])) continue;
            REAL d[3];
            REAL r2 = FN(geom)(pi, pc, j, sp[k].sh, d);
            if (!(r2 < H2)) continue;
            if (own && j == i) continue;
            ++cnt;
            REAL invR = (REAL)0, rr = (REAL)0;
            if (r2 > (REAL)0) {
                invR = (REAL)1 / (REAL)sqrt(r2);
                rr = r2 * invR;
            }
            REAL q = rr * invh;
            REAL t = m[j] * FN(kshape)(cfg->kernel, q);
            acc += t;
            sc += fabs((double)t);
        }
    }
    REAL self = m[i] * f0;
    rho[i] = wfac * (acc + self);
    if (st->nnb) st->nnb[i] = cnt;
    if (st->sc_rho) st->sc_rho[i] = fabs((double)wfac) * (sc + fabs((double)self));
}

/* Cleaning speed after a density pass: max fast speed over targets (new; Tricco & Price 2012
 * sec. 2); 0 without cleaning. */
static double FN(clean_speed)(const sph_run_config* cfg, const orc_state* st)
{
    double chmax = 0.0;
    if (cfg->mhd && cfg->cleaning) {
        int64_t n = st->n;
        FN(prim)* pr = (FN(prim)*)malloc((size_t)n * sizeof(FN(prim)) + 8);
        FN(prims)(cfg, st, pr);
        for (int64_t i = 0; i < n; ++i)
            if (is_target(st->species[i], st->mask[i]) && (double)pr[i].c > chmax) chmax = (double)pr[i].c;
        free(pr);
    }
    return chmax;
}

/* Density summation over the supercell stencil (DensitySummation.hpp:50-138 semantics). */
int FN(orc_density)(const sph_run_config* cfg, int32_t prec, orc_state* st, const uint32_t* cell_start)
{
    orc_grid g;
    if (orc_make_grid(cfg, &g)) return -1;
    int64_t n = st->n;
    REAL* pc[3];
    for (int a = 0; a < 3; ++a) pc[a] = (REAL*)malloc((size_t)n * sizeof(REAL) + 8);
    FN(pair_coords)(&g, prec, st, pc);
    const REAL H2 = (prec == ORC_PREC_F64) ? (REAL)g.H2d : (REAL)g.H2;
    const REAL f0 = FN(kshape)(cfg->kernel, (REAL)0);

#pragma omp parallel for schedule(dynamic, 16)
    for (int cell = 0; cell < g.ncell; ++cell) {
        FN(span) sp[125];
        const int nsp = FN(cell_spans)(&g, prec, cfg, cell_start, cell, sp);
        for (int64_t i = cell_start[cell]; i < (int64_t)cell_start[cell + 1]; ++i) {
            if (!is_target(st->species[i], st->mask[i])) {
                if (st->nnb) st->nnb[i] = 0;
                continue;
            }
            FN(density_one)(cfg, st, pc, H2, f0, i, sp, nsp);
        }
    }
    for (int a = 0; a < 3; ++a) free(pc[a]);
    st->c_h = FN(clean_speed)(cfg, st);
    return 0;
}

/* Inputs of the force pass shared by all targets. */
typedef struct FN(fctx) {
    REAL* pc[3];
    const FN(prim)* pr;
    REAL H2, alpha, beta, borve, ch, csig;
    int kern, mhd, clean;
} FN(fctx);

static void FN(fctx_init)(const sph_run_config* cfg, const orc_state* st, REAL H2, REAL* const pc[3],
                          const FN(prim)* pr, FN(fctx)* F)
{
    for (int a = 0; a < 3; ++a) F->pc[a] = pc[a];
    F->pr = pr;
    F->H2 = H2;
    F->kern = cfg->kernel;
    F->mhd = cfg->mhd;
    F->clean = cfg->mhd && cfg->cleaning;
    F->alpha = (REAL)cfg->alpha_visc;
    F->beta = (REAL)cfg->beta_visc;
    F->borve = (cfg->mhd && cfg->borve) ? (REAL)1 : (REAL)0;
    F->ch = (REAL)st->c_h;
    F->csig = (REAL)cfg->clean_sigma;
}

/* Forces on target i over the source spans: HydroInteraction (HydroForces.hpp:139-175) + Monaghan
 * viscosity + SPMHD + cleaning. */
static void FN(force_one)(const FN(fctx)* F, orc_state* st, int64_t i, const FN(span)* sp, int nsp)
{
    REAL* const* pc = F->pc;
    const FN(prim)* pr = F->pr;
    const REAL H2 = F->H2;
    const int kern = F->kern, mhd = F->mhd, clean = F->clean;
    const REAL alpha = F->alpha, beta = F->beta, borve = F->borve, ch = F->ch, csig = F->csig;
    const REAL* m = RA(st->mass);
    const REAL* hh = RA(st->h);
    const REAL* rho = RA(st->rho);
    const REAL *vx = RA(st->vel[0]), *vy = RA(st->vel[1]), *vz = RA(st->vel[2]);
    const REAL *Bx = mhd ? RA(st->B[0]) : NULL, *By = mhd ? RA(st->B[1]) : NULL,
               *Bz = mhd ? RA(st->B[2]) : NULL;
    const FN(prim) P_i = pr[i];
    REAL pi[3] = {pc[0][i], pc[1][i], pc[2][i]};
    REAL bi[3] = {0, 0, 0};
    if (mhd) { bi[0] = Bx[i]; bi[1] = By[i]; bi[2] = Bz[i]; }
    REAL ax = 0, ay = 0, az = 0, DV = 0, AVU = 0, vmax = 0;
    REAL Ti = 0, Tj = 0, TBx = 0, TBy = 0, TBz = 0, VBx = 0, VBy = 0, VBz = 0, DB = 0;
    REAL PSx = 0, PSy = 0, PSz = 0;
    double sa = 0, sdu = 0, sdb = 0, sdiv = 0, sdvv = 0;
    for (int k = 0; k < nsp; ++k) {
        const int own = sp[k].own;
        for (int64_t j = sp[k].b; j < sp[k].e; ++j) {
            if (!is_source(st->species[j], st->mask[j])) continue;
            REAL d[3];
            REAL r2 = FN(geom)(pi, pc, j, sp[k].sh, d);
            if (!(r2 < H2)) continue;
            if (own && j == i) continue;
            const FN(prim) P_j = pr[j];
            REAL invR = (REAL)0, rr = (REAL)0;
            if (r2 > (REAL)0) {
                invR = (REAL)1 / (REAL)sqrt(r2);
                rr = r2 * invR;
            }
            REAL s_i = FN(kgrad)(kern, rr, invR, rr * P_i.invh, P_i.gfac);
            REAL s_j = FN(kgrad)(kern, rr, invR, rr * P_j.invh, P_j.gfac);
            REAL dvx = vx[i] - vx[j], dvy = vy[i] - vy[j], dvz = vz[i] - vz[j];
            REAL vr = dvx * d[0] + dvy * d[1] + dvz * d[2];
            REAL ms_i = m[j] * s_i, ms_j = m[j] * s_j;
            /* pressure (+ magnetic pressure) gradient, symmetric form */
            REAL fiso = P_i.pr * ms_i + P_j.pr * ms_j;
            REAL cax = -fiso * d[0], cay = -fiso * d[1], caz = -fiso * d[2];
            DV += ms_i * vr;
            /* scales use the magnitudes before cancellation: sum_a |dv_a d_a| for v.r (and
             * sum_a |B_a d_a| for B.r below), the bound of an FP32 evaluation in any order */
            const double vdn = fabs((double)dvx * (double)d[0]) + fabs((double)dvy * (double)d[1])
                               + fabs((double)dvz * (double)d[2]);
            sdvv += fabs((double)ms_i) * vdn;
            sdu += fabs((double)(P_i.ph * ms_i)) * vdn;
            /* signal velocity */
            REAL w = vr * invR;
            REAL vs = P_i.c + P_j.c - beta * (w < (REAL)0 ? w : (REAL)0);
            if (vs > vmax) vmax = vs;
            /* artificial viscosity */
            if (alpha > (REAL)0 && vr < (REAL)0) {
                REAL hb = (REAL)0.5 * (hh[i] + hh[j]);
                REAL mu = hb * vr / (r2 + (REAL)0.01 * hb * hb);
                REAL cb = (REAL)0.5 * (P_i.c + P_j.c);
                REAL rb = (REAL)0.5 * (rho[i] + rho[j]);
                REAL Pi = (beta * mu - alpha * cb) * mu / rb;
                REAL gv = Pi * (REAL)0.5 * (ms_i + ms_j);
                cax -= gv * d[0];
                cay -= gv * d[1];
                caz -= gv * d[2];
                AVU += gv * vr;
                sdu += fabs((double)gv) * vdn;
            }
            ax += cax;
            ay += cay;
            az += caz;
            double cnorm = sqrt((double)cax * (double)cax + (double)cay * (double)cay + (double)caz * (double)caz);
            if (mhd) {
                REAL Bir = bi[0] * d[0] + bi[1] * d[1] + bi[2] * d[2];
                REAL Bjr = Bx[j] * d[0] + By[j] * d[1] + Bz[j] * d[2];
                REAL ti = Bir * P_i.bdr * ms_i;
                REAL tj = Bjr * P_j.bdr * ms_j;
                Ti += ti;
                Tj += tj;
                TBx += tj * Bx[j];
                TBy += tj * By[j];
                TBz += tj * Bz[j];
                REAL wv = ms_i * Bir;
                VBx += wv * dvx;
                VBy += wv * dvy;
                VBz += wv * dvz;
                REAL db = ms_i * (Bir - Bjr);
                DB += db;
                double bn_i = sqrt((double)bi[0] * (double)bi[0] + (double)bi[1] * (double)bi[1] + (double)bi[2] * (double)bi[2]);
                double bn_j = sqrt((double)Bx[j] * (double)Bx[j] + (double)By[j] * (double)By[j] + (double)Bz[j] * (double)Bz[j]);
                const double bdi = fabs((double)bi[0] * (double)d[0]) + fabs((double)bi[1] * (double)d[1])
                                   + fabs((double)bi[2] * (double)d[2]);
                const double bdj = fabs((double)Bx[j] * (double)d[0]) + fabs((double)By[j] * (double)d[1])
                                   + fabs((double)Bz[j] * (double)d[2]);
                const double ati = fabs((double)(P_i.bdr * ms_i)) * bdi;
                const double atj = fabs((double)(P_j.bdr * ms_j)) * bdj;
                cnorm += ati * bn_i + atj * bn_j + (double)borve * (ati + atj) * bn_i;
                double dvn = sqrt((double)dvx * (double)dvx + (double)dvy * (double)dvy + (double)dvz * (double)dvz);
                sdb += (fabs((double)ms_i) * bdi * dvn + bn_i * fabs((double)ms_i) * vdn) / (double)rho[i];
                sdiv += fabs((double)ms_i) * (bdi + bdj) / (double)rho[i];
                if (clean) {
                    REAL fps = P_i.psr * ms_i + P_j.psr * ms_j;
                    PSx += fps * d[0];
                    PSy += fps * d[1];
                    PSz += fps * d[2];
                    double dn = sqrt((double)d[0] * (double)d[0] + (double)d[1] * (double)d[1] + (double)d[2] * (double)d[2]);
                    sdb += (double)rho[i] * (fabs((double)(P_i.psr * ms_i)) + fabs((double)(P_j.psr * ms_j))) * dn;
                }
            }
            sa += cnorm;
        }
    }
    REAL invrho = (REAL)1 / rho[i];
    if (mhd) {
        REAL tcoef = Ti - borve * (Ti + Tj);
        ax = ax + TBx + bi[0] * tcoef;
        ay = ay + TBy + bi[1] * tcoef;
        az = az + TBz + bi[2] * tcoef;
    }
    RA(st->dvdt[0])[i] = ax;
    RA(st->dvdt[1])[i] = ay;
    RA(st->dvdt[2])[i] = az;
    RA(st->dudt)[i] = P_i.ph * DV + (REAL)0.5 * AVU;
    REAL vsg = (REAL)2 * P_i.c;
    if (vmax > vsg) vsg = vmax;
    RA(st->vsig)[i] = vsg;
    if (mhd) {
        REAL divB = -DB * invrho;
        REAL divv = -DV * invrho;
        REAL dbx = -(VBx - bi[0] * DV) * invrho;
        REAL dby = -(VBy - bi[1] * DV) * invrho;
        REAL dbz = -(VBz - bi[2] * DV) * invrho;
        REAL dpsi = (REAL)0;
        if (clean) {
            dbx -= rho[i] * PSx;
            dby -= rho[i] * PSy;
            dbz -= rho[i] * PSz;
            REAL psi_i = RA(st->psi)[i];
            dpsi = -ch * ch * divB - psi_i * csig * ch * P_i.invh - (REAL)0.5 * psi_i * divv;
        }
        RA(st->dbdt[0])[i] = dbx;
        RA(st->dbdt[1])[i] = dby;
        RA(st->dbdt[2])[i] = dbz;
        RA(st->dpsidt)[i] = dpsi;
        RA(st->divb)[i] = divB;
    }
    if (st->sc_a) st->sc_a[i] = sa;
    if (st->sc_du) st->sc_du[i] = sdu;
    if (st->sc_db) st->sc_db[i] = sdb;
    if (st->sc_divb) st->sc_divb[i] = sdiv;
    if (st->sc_dpsi && clean) {
        double psi_i = (double)RA(st->psi)[i];
        st->sc_dpsi[i] = (double)ch * (double)ch * sdiv
                         + fabs(psi_i * (double)csig * (double)ch * (double)P_i.invh)
                         + 0.5 * fabs(psi_i) * sdvv / (double)rho[i];
    }
}

/* Force pass over the supercell stencil. */
int FN(orc_forces)(const sph_run_config* cfg, int32_t prec, orc_state* st, const uint32_t* cell_start)
{
    orc_grid g;
    if (orc_make_grid(cfg, &g)) return -1;
    int64_t n = st->n;
    REAL* pc[3];
    for (int a = 0; a < 3; ++a) pc[a] = (REAL*)malloc((size_t)n * sizeof(REAL) + 8);
    FN(pair_coords)(&g, prec, st, pc);
    FN(prim)* pr = (FN(prim)*)malloc((size_t)n * sizeof(FN(prim)) + 8);
    FN(prims)(cfg, st, pr);
    FN(fctx) F;
    FN(fctx_init)(cfg, st, (prec == ORC_PREC_F64) ? (REAL)g.H2d : (REAL)g.H2, pc, pr, &F);

#pragma omp parallel for schedule(dynamic, 16)
    for (int cell = 0; cell < g.ncell; ++cell) {
        FN(span) sp[125];
        const int nsp = FN(cell_spans)(&g, prec, cfg, cell_start, cell, sp);
        for (int64_t i = cell_start[cell]; i < (int64_t)cell_start[cell + 1]; ++i) {
            if (!is_target(st->species[i], st->mask[i])) continue;
            FN(force_one)(&F, st, i, sp, nsp);
        }
    }
    free(pr);
    for (int a = 0; a < 3; ++a) free(pc[a]);
    return 0;
}

/* Drift: PushVelocity (PushVelocity.hpp:35-42) for Movable species, then the periodic wrap (skipped
 * when wrap = 0: the region traversal keeps positions continuous, as the reference does). */
static int FN(drift_impl)(const sph_run_config* cfg, int32_t prec, orc_state* st, double dt, int wrap)
{
    orc_grid g;
    if (orc_make_grid(cfg, &g)) return -1;
    int64_t n = st->n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        if (!is_movable(st->species[i], st->mask[i])) continue;
        for (int a = 0; a < 3; ++a) {
            REAL v = RA(st->vel[a])[i];
            if (prec == ORC_PREC_PLAIN) {
                float x = (float)st->pos[a][i];
                float dts = (float)dt;
                float vv = (float)v;
                x = x + vv * dts;
                if (wrap && cfg->periodic[a]) {
                    float hi = g.lo32[a] + g.L32[a];
                    if (x >= hi) x = x - g.L32[a];
                    else if (x < g.lo32[a]) x = x + g.L32[a];
                }
                st->pos[a][i] = (double)x;
            } else {
                double x = st->pos[a][i];
                x = x + (double)v * dt;
                if (wrap && cfg->periodic[a]) {
                    double hi = g.lo[a] + g.L[a];
                    if (x >= hi) x = x - g.L[a];
                    else if (x < g.lo[a]) x = x + g.L[a];
                }
                st->pos[a][i] = x;
            }
        }
    }
    return 0;
}

int FN(orc_drift)(const sph_run_config* cfg, int32_t prec, orc_state* st, double dt)
{
    return FN(drift_impl)(cfg, prec, st, dt, 1);
}

/* CFL dt of one target (new; Monaghan 1997 signal velocity, Courant + force criteria). */
static inline double FN(dt_particle)(const sph_run_config* cfg, int32_t prec, REAL h, REAL vsig, REAL ax,
                                     REAL ay, REAL az)
{
    if (prec == ORC_PREC_PLAIN) {
        float dti = (float)cfg->cfl * (float)h / (float)vsig;
        float a2 = (float)ax * (float)ax + (float)ay * (float)ay + (float)az * (float)az;
        if (a2 > 0.0f) {
            float dtf = (float)cfg->cfl_force * sqrtf((float)h / sqrtf(a2));
            if (dtf < dti) dti = dtf;
        }
        return (double)dti;
    }
    double dti = (double)cfg->cfl * (double)h / (double)vsig;
    double a2 = (double)ax * (double)ax + (double)ay * (double)ay + (double)az * (double)az;
    if (a2 > 0.0) {
        double dtf = (double)cfg->cfl_force * sqrt((double)h / sqrt(a2));
        if (dtf < dti) dti = dtf;
    }
    return dti;
}

/* FP64 reductions over targets: min dt, mass, momentum, energy. */
int FN(orc_reduce)(const sph_run_config* cfg, int32_t prec, orc_state* st, sph_conserved* out)
{
    double M = 0, Px = 0, Py = 0, Pz = 0, E = 0, dtmin = INFINITY;
    for (int64_t i = 0; i < st->n; ++i) {
        if (!is_target(st->species[i], st->mask[i])) continue;
        double m = (double)RA(st->mass)[i];
        double vx = (double)RA(st->vel[0])[i], vy = (double)RA(st->vel[1])[i], vz = (double)RA(st->vel[2])[i];
        double v2 = vx * vx + vy * vy + vz * vz;
        double b2 = 0;
        if (cfg->mhd) {
            double bx = (double)RA(st->B[0])[i], by = (double)RA(st->B[1])[i], bz = (double)RA(st->B[2])[i];
            b2 = bx * bx + by * by + bz * bz;
        }
        M += m;
        Px += m * vx;
        Py += m * vy;
        Pz += m * vz;
        E += m * (0.5 * v2 + (double)RA(st->u)[i] + 0.5 * b2 / (double)RA(st->rho)[i]);
        double dti = FN(dt_particle)(cfg, prec, RA(st->h)[i], RA(st->vsig)[i], RA(st->dvdt[0])[i],
                                     RA(st->dvdt[1])[i], RA(st->dvdt[2])[i]);
        if (dti < dtmin) dtmin = dti;
    }
    if (out) {
        out->mass = M;
        out->momentum[0] = Px;
        out->momentum[1] = Py;
        out->momentum[2] = Pz;
        out->energy = E;
        out->dt_next = cfg->dt_fixed > 0 ? cfg->dt_fixed : dtmin;
        out->time = st->time;
        out->c_h = st->c_h;
    }
    return 0;
}

/* Kick: EulerIntegrate (EulerIntegrate.hpp:43-52) on Movable+Thermodynamic, extended to B and psi,
 * then the reductions of the end-of-step state. */
int FN(orc_kick)(const sph_run_config* cfg, int32_t prec, orc_state* st, double dt, sph_conserved* out)
{
    REAL dtr = (REAL)dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < st->n; ++i) {
        if (!is_target(st->species[i], st->mask[i])) continue;
        for (int a = 0; a < 3; ++a) RA(st->vel[a])[i] = RA(st->vel[a])[i] + RA(st->dvdt[a])[i] * dtr;
        RA(st->u)[i] = RA(st->u)[i] + RA(st->dudt)[i] * dtr;
        if (cfg->mhd) {
            for (int a = 0; a < 3; ++a) RA(st->B[a])[i] = RA(st->B[a])[i] + RA(st->dbdt[a])[i] * dtr;
            RA(st->psi)[i] = RA(st->psi)[i] + RA(st->dpsidt)[i] * dtr;
        }
    }
    st->time += dt;
    return FN(orc_reduce)(cfg, prec, st, out);
}

int FN(orc_prime)(const sph_run_config* cfg, int32_t prec, orc_state* st, sph_conserved* out)
{
    orc_grid g;
    if (orc_make_grid(cfg, &g)) return -1;
    uint32_t* cs = (uint32_t*)malloc(((size_t)g.ncell + 1) * 4);
    int rc = FN(orc_resort)(cfg, prec, st, cs);
    if (!rc) rc = FN(orc_density)(cfg, prec, st, cs);
    if (!rc) rc = FN(orc_forces)(cfg, prec, st, cs);
    sph_conserved c;
    if (!rc) rc = FN(orc_reduce)(cfg, prec, st, &c);
    if (!rc) st->dt = c.dt_next;
    if (out) *out = c;
    free(cs);
    return rc;
}

/* Simulation::runOneStep order (Simulation.x.cpp:196-310): drift, resort, density, forces, kick. */
int FN(orc_step)(const sph_run_config* cfg, int32_t prec, orc_state* st, int32_t n_steps, sph_conserved* out)
{
    orc_grid g;
    if (orc_make_grid(cfg, &g)) return -1;
    uint32_t* cs = (uint32_t*)malloc(((size_t)g.ncell + 1) * 4);
    int rc = 0;
    sph_conserved c;
    memset(&c, 0, sizeof(c));
    for (int s = 0; s < n_steps && !rc; ++s) {
        double dt = cfg->dt_fixed > 0 ? cfg->dt_fixed : st->dt;
        if (!(dt > 0)) { rc = -4; break; }
        /* dt = +inf: no live target bounds the CFL step, so the run does not advance (no drift, no
         * kick, no time); its derivatives and sums are still evaluated */
        const int advance = dt < HUGE_VAL;
        if (advance) rc = FN(orc_drift)(cfg, prec, st, dt);
        if (!rc) rc = FN(orc_resort)(cfg, prec, st, cs);
        if (!rc) rc = FN(orc_density)(cfg, prec, st, cs);
        if (!rc) rc = FN(orc_forces)(cfg, prec, st, cs);
        if (!rc) rc = advance ? FN(orc_kick)(cfg, prec, st, dt, &c) : FN(orc_reduce)(cfg, prec, st, &c);
        if (!rc) st->dt = c.dt_next;
    }
    if (out) *out = c;
    free(cs);
    return rc;
}

/* Brute-force O(N^2) neighbour counts over all periodic images (absolute coordinates only):
 * the check on the supercell stencil. */
int FN(orc_brute_counts)(const sph_run_config* cfg, int32_t prec, orc_state* st)
{
    if (prec == ORC_PREC_MIXED) return -1;
    orc_grid g;
    if (orc_make_grid(cfg, &g)) return -1;
    int64_t n = st->n;
    REAL* pc[3];
    for (int a = 0; a < 3; ++a) pc[a] = (REAL*)malloc((size_t)n * sizeof(REAL) + 8);
    FN(pair_coords)(&g, prec, st, pc);
    const REAL H2 = (prec == ORC_PREC_F64) ? (REAL)g.H2d : (REAL)g.H2;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; ++i) {
        if (!is_target(st->species[i], st->mask[i])) { st->nnb[i] = 0; continue; }
        uint32_t cnt = 0;
        REAL pi[3] = {pc[0][i], pc[1][i], pc[2][i]};
        for (int64_t j = 0; j < n; ++j) {
            if (!is_source(st->species[j], st->mask[j])) continue;
            for (int wz = -1; wz <= 1; ++wz)
                for (int wy = -1; wy <= 1; ++wy)
                    for (int wx = -1; wx <= 1; ++wx) {
                        int w[3] = {wx, wy, wz};
                        REAL sh[3];
                        int ok = 1;
                        for (int a = 0; a < 3; ++a) {
                            if (!cfg->periodic[a] && w[a] != 0) ok = 0;
                            sh[a] = (prec == ORC_PREC_PLAIN) ? (REAL)w[a] * (REAL)g.L32[a]
                                                             : (REAL)((double)w[a] * g.L[a]);
                        }
                        if (!ok) continue;
                        if (j == i && wx == 0 && wy == 0 && wz == 0) continue;
                        REAL d[3];
                        REAL r2 = FN(geom)(pi, pc, j, sh, d);
                        if (r2 < H2) ++cnt;
                    }
        }
        st->nnb[i] = cnt;
    }
    for (int a = 0; a < 3; ++a) free(pc[a]);
    return 0;
}

/* ---------------------------------------------------------------------------------------------
 * Region traversal: the reference's own neighbour algorithm, restated for the CPU baseline. The
 * state is laid out by orc_regions_layout: regions (blocks of supercells at set-up) keep their
 * particles for the whole run, in 64-slot frames (ParticleRegion.hpp:37-62, numFrameSlots = 64 in
 * param/memory.param:35), unused slots dead (mask 0). Each step:
 *   - refits each region's AABB to its live sources (UpdateRegionBounds, RegionBoundsUpdate.hpp:39-165);
 *   - brute-force region search: every source region whose AABB intersects the target region's AABB
 *     expanded by H, counted then written as CSR (FindNeighbourRegionsFunctor,
 *     NeighbourRegions.hpp:55-135; calculateNeighbourEntrySender :140-230);
 *   - density and force passes frame by frame: every target slot against every slot of every frame
 *     of every neighbour region, dead slots tested and rejected (interactWithNeighbourFrame,
 *     ParticleParticleInteraction.hpp:219-314), one OpenMP iteration per target frame.
 * Periodic axes (which the reference lacks) test each source AABB under the images -L, 0, +L and
 * keep one CSR entry per overlapping image; positions are not wrapped, so a region stays compact.
 * The expansion is H widened by 1e-6 relative, so FP32 rounding of the bounds never drops a pair
 * within H. No precision mode but absolute coordinates (plain FP32, FP64).
 * --------------------------------------------------------------------------------------------- */

/* Images (bit w+1 for shift w L, w in -1..1) under which source interval [s0, s1] meets [t0, t1];
 * the inclusive test of AABB.hpp:116-119. */
static inline unsigned FN(axis_images)(REAL t0, REAL t1, REAL s0, REAL s1, REAL L, int periodic)
{
    unsigned m = !(t0 > s1 || t1 < s0) ? 2u : 0u;
    if (periodic) {
        if (!(t0 > s1 - L || t1 < s0 - L)) m |= 1u;
        if (!(t0 > s1 + L || t1 < s0 + L)) m |= 4u;
    }
    return m;
}

static const int FN(pop3)[8] = {0, 1, 1, 2, 1, 2, 2, 3};

/* Brute-force region search (FindNeighbourRegionsFunctor, NeighbourRegions.hpp:55-135): for each
 * target box, every source box that meets it expanded by `expand`, once per periodic image that
 * does. Boxes are SoA: lo x, y, z then hi x, y, z, each an array of the box count.
 * Count pass (ent == NULL): off[r + 1] = entries of target r, then scanned with off[0] = 0 (the
 * Count kernel and inclusiveScanHost, NeighbourRegions.hpp:167-193). Write pass: target r's entries
 * from off[r], sources in index order, images in (z, y, x) order (the reference orders a row by
 * block atomics, so there only the set of a row is fixed). Returns the entry count. */
static int64_t FN(region_search)(int32_t nt, const REAL* tb, int32_t ns, const REAL* sb, REAL expand,
                                 const REAL L[3], const int32_t periodic[3], int64_t* off, int32_t* ent,
                                 uint8_t* img)
{
    const REAL* slo[3] = {sb, sb + ns, sb + 2 * (size_t)ns};
    const REAL* shi[3] = {sb + 3 * (size_t)ns, sb + 4 * (size_t)ns, sb + 5 * (size_t)ns};
#pragma omp parallel for schedule(dynamic, 64)
    for (int32_t r = 0; r < nt; ++r) {
        REAL t0[3], t1[3];
        for (int a = 0; a < 3; ++a) {
            t0[a] = tb[(size_t)a * nt + r] - expand;
            t1[a] = tb[(size_t)(3 + a) * nt + r] + expand;
        }
        int64_t w = ent ? off[r] : 0;
        for (int32_t q = 0; q < ns; ++q) {
            unsigned mx = FN(axis_images)(t0[0], t1[0], slo[0][q], shi[0][q], L[0], periodic[0]);
            if (!mx) continue;
            unsigned my = FN(axis_images)(t0[1], t1[1], slo[1][q], shi[1][q], L[1], periodic[1]);
            if (!my) continue;
            unsigned mz = FN(axis_images)(t0[2], t1[2], slo[2][q], shi[2][q], L[2], periodic[2]);
            if (!ent) {
                w += FN(pop3)[mx] * FN(pop3)[my] * FN(pop3)[mz];
                continue;
            }
            for (int wz = 0; wz < 3; ++wz)
                for (int wy = 0; wy < 3; ++wy)
                    for (int wx = 0; wx < 3; ++wx)
                        if (((mx >> wx) & 1u) && ((my >> wy) & 1u) && ((mz >> wz) & 1u)) {
                            ent[w] = q;
                            img[w] = (uint8_t)(wx + 3 * wy + 9 * wz);
                            ++w;
                        }
        }
        if (!ent) off[r + 1] = w;
    }
    if (!ent) {
        off[0] = 0;
        for (int32_t r = 0; r < nt; ++r) off[r + 1] += off[r];
    }
    return off[nt];
}

/* The search on caller-given boxes (REAL arrays, 6 per box as above). */
int64_t FN(orc_region_search)(int32_t nt, const void* tbox, int32_t ns, const void* sbox, double expand,
                              const double L[3], const int32_t periodic[3], int64_t* off, int32_t* ent, uint8_t* img)
{
    REAL Lr[3];
    for (int a = 0; a < 3; ++a) Lr[a] = (REAL)L[a];
    return FN(region_search)(nt, (const REAL*)tbox, ns, (const REAL*)sbox, (REAL)expand, Lr, periodic, off, ent,
                             img);
}

/* AABB refit (UpdateRegionBounds, RegionBoundsUpdate.hpp:39-165): each region's box is the min / max
 * of its live sources' pair coordinates; a region without any gets the empty box (+inf, -inf), which
 * meets nothing in the search. Boxes in the search's SoA layout. */
static void FN(region_refit)(const orc_state* st, REAL* const pc[3], int32_t R, const int64_t* region_start, REAL* bb)
{
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r < R; ++r) {
        REAL l[3], h[3];
        for (int a = 0; a < 3; ++a) {
            l[a] = (REAL)HUGE_VAL;
            h[a] = (REAL)-HUGE_VAL;
        }
        for (int64_t i = region_start[r]; i < region_start[r + 1]; ++i) {
            if (!is_source(st->species[i], st->mask[i])) continue;
            for (int a = 0; a < 3; ++a) {
                if (pc[a][i] < l[a]) l[a] = pc[a][i];
                if (pc[a][i] > h[a]) h[a] = pc[a][i];
            }
        }
        for (int a = 0; a < 3; ++a) {
            bb[(size_t)a * R + r] = l[a];
            bb[(size_t)(3 + a) * R + r] = h[a];
        }
    }
}

/* The refit on a state laid out in regions (absolute coordinates: plain FP32 or FP64). */
int FN(orc_regions_refit)(const sph_run_config* cfg, int32_t prec, const orc_state* st, int32_t n_regions,
                          const int64_t* region_start, void* box)
{
    if (prec == ORC_PREC_MIXED || n_regions < 0) return -1;
    orc_grid g;
    if (orc_make_grid(cfg, &g)) return -1;
    REAL* pc[3];
    for (int a = 0; a < 3; ++a) pc[a] = (REAL*)malloc((size_t)st->n * sizeof(REAL) + 8);
    if (!pc[0] || !pc[1] || !pc[2]) return -2;
    FN(pair_coords)(&g, prec, st, pc);
    FN(region_refit)(st, pc, n_regions, region_start, (REAL*)box);
    for (int a = 0; a < 3; ++a) free(pc[a]);
    return 0;
}

/* The CSR row of region r as spans: each neighbour region's slots under its image shift. */
static int FN(region_spans)(const int64_t* region_start, const int64_t* off, const int32_t* ent, const uint8_t* img,
                            int32_t r, const REAL Lr[3], FN(span)* sp)
{
    int nsp = 0;
    for (int64_t e = off[r]; e < off[r + 1]; ++e) {
        const int q = ent[e], code = img[e];
        const int w[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
        sp[nsp].b = region_start[q];
        sp[nsp].e = region_start[q + 1];
        sp[nsp].own = (q == r && code == 13);
        for (int a = 0; a < 3; ++a) sp[nsp].sh[a] = (REAL)w[a] * Lr[a];
        ++nsp;
    }
    return nsp;
}

int FN(orc_regions_step)(const sph_run_config* cfg, int32_t prec, orc_state* st, int32_t n_regions,
                         const int64_t* region_start, int32_t n_steps, int64_t* csr_entries, sph_conserved* out)
{
    if (prec == ORC_PREC_MIXED || n_regions < 1 || n_steps < 0) return -1;
    orc_grid g;
    if (orc_make_grid(cfg, &g)) return -1;
    const int64_t n = st->n;
    if (n % 64 != 0 || region_start[0] != 0 || region_start[n_regions] != n) return -1;
    for (int32_t r = 0; r < n_regions; ++r)
        if (region_start[r + 1] <= region_start[r] || region_start[r + 1] % 64 != 0) return -1;
    const int64_t nframes = n / 64;
    const int32_t R = n_regions;
    const REAL H = (prec == ORC_PREC_F64) ? (REAL)g.Hd : (REAL)g.H;
    const REAL H2 = (prec == ORC_PREC_F64) ? (REAL)g.H2d : (REAL)g.H2;
    const REAL Hs = H + H * (REAL)1e-6;
    REAL Lr[3];
    for (int a = 0; a < 3; ++a) Lr[a] = (prec == ORC_PREC_F64) ? (REAL)g.L[a] : (REAL)g.L32[a];
    const REAL f0 = FN(kshape)(cfg->kernel, (REAL)0);

    int32_t* frame_region = (int32_t*)malloc((size_t)nframes * 4 + 4);
    int64_t* off = (int64_t*)malloc(((size_t)R + 1) * 8);
    REAL* bb = (REAL*)malloc((size_t)R * 6 * sizeof(REAL));
    REAL* pc[3];
    for (int a = 0; a < 3; ++a) pc[a] = (REAL*)malloc((size_t)n * sizeof(REAL) + 8);
    FN(prim)* pr = (FN(prim)*)malloc((size_t)n * sizeof(FN(prim)) + 8);
    int rc = (frame_region && off && bb && pc[0] && pc[1] && pc[2] && pr) ? 0 : -2;
    if (!rc)
        for (int32_t r = 0; r < R; ++r)
            for (int64_t f = region_start[r] / 64; f < region_start[r + 1] / 64; ++f) frame_region[f] = r;
    /* span buffers, one per thread, sized to the longest CSR row of each step */
    FN(span)* spans = NULL;
    const int nthr = orc_num_threads();
    sph_conserved c;
    memset(&c, 0, sizeof(c));
    const int prime = (n_steps == 0);
    for (int s = 0; s < (prime ? 1 : n_steps) && !rc; ++s) {
        double dt = cfg->dt_fixed > 0 ? cfg->dt_fixed : st->dt;
        if (!prime && !(dt > 0)) { rc = -4; break; }
        const int advance = !prime && dt < HUGE_VAL;
        if (advance) rc = FN(drift_impl)(cfg, prec, st, dt, 0);
        if (rc) break;
        FN(pair_coords)(&g, prec, st, pc);

        FN(region_refit)(st, pc, R, region_start, bb);

        /* region search: Count, scan, Write (every region is both a target and a source) */
        const int64_t total = FN(region_search)(R, bb, R, bb, Hs, Lr, cfg->periodic, off, NULL, NULL);
        int64_t maxrow = 0;
        for (int32_t r = 0; r < R; ++r)
            if (off[r + 1] - off[r] > maxrow) maxrow = off[r + 1] - off[r];
        if (csr_entries) *csr_entries = total;
        int32_t* ent = (int32_t*)malloc((size_t)total * 4 + 4);
        uint8_t* img = (uint8_t*)malloc((size_t)total + 1);
        free(spans);
        spans = (FN(span)*)malloc((size_t)nthr * (size_t)(maxrow + 1) * sizeof(FN(span)));
        if (!ent || !img || !spans) {
            free(ent);
            free(img);
            rc = -2;
            break;
        }
        FN(region_search)(R, bb, R, bb, Hs, Lr, cfg->periodic, off, ent, img);

        /* density pass, one target frame per iteration */
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t f = 0; f < nframes; ++f) {
            FN(span)* sp = spans + (size_t)orc_thread_num() * (size_t)(maxrow + 1);
            const int nsp = FN(region_spans)(region_start, off, ent, img, frame_region[f], Lr, sp);
            for (int64_t i = f * 64; i < f * 64 + 64; ++i) {
                if (!is_target(st->species[i], st->mask[i])) {
                    if (st->nnb) st->nnb[i] = 0;
                    continue;
                }
                FN(density_one)(cfg, st, pc, H2, f0, i, sp, nsp);
            }
        }
        st->c_h = FN(clean_speed)(cfg, st);

        /* force pass, same traversal */
        FN(prims)(cfg, st, pr);
        FN(fctx) F;
        FN(fctx_init)(cfg, st, H2, pc, pr, &F);
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t f = 0; f < nframes; ++f) {
            FN(span)* sp = spans + (size_t)orc_thread_num() * (size_t)(maxrow + 1);
            const int nsp = FN(region_spans)(region_start, off, ent, img, frame_region[f], Lr, sp);
            for (int64_t i = f * 64; i < f * 64 + 64; ++i)
                if (is_target(st->species[i], st->mask[i])) FN(force_one)(&F, st, i, sp, nsp);
        }
        free(ent);
        free(img);

        rc = advance ? FN(orc_kick)(cfg, prec, st, dt, &c) : FN(orc_reduce)(cfg, prec, st, &c);
        if (!rc) st->dt = c.dt_next;
    }
    if (out) *out = c;
    free(spans);
    free(frame_region);
    free(off);
    free(bb);
    for (int a = 0; a < 3; ++a) free(pc[a]);
    free(pr);
    return rc;
}

#undef FN
#undef RA
