/*
 * sph_oracle.c -- CPU oracle for the SPH(-MHD) particle step. TEST INFRASTRUCTURE ONLY: loaded by
 * tests, the smoke check and the benchmark's CPU-baseline leg; never by the product library.
 *
 * Compile with -ffp-contract=off so that every a*b+c rounds twice, as the restated reference code
 * does on its CPU backends; the binning and neighbour-count contract with the HIP path depends on it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "sph_oracle.h"

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_thread_num(void)
{
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

/* Effective supercell factor: <= 0 (unset) is 1; below 1/2 it is 1/2. Edges of at least H reach the
 * 27-cell stencil, edges of at least H/2 the 125-cell stencil. */
static double cell_factor_of(float cf)
{
    if (!(cf > 0.0f)) return 1.0;
    return cf < 0.5f ? 0.5 : (double)cf;
}

/* Uniform supercell grid: per axis the largest cell count whose edge is >= cell_factor * H with a
 * 1e-4 relative margin (so FP32 rounding of the cell index can never drop a pair within H).
 * This replaces the reference's irregular region AABBs and their brute-force overlap search
 * (packages/pmacc/include/spmacc/particles/regions/NeighbourRegions.hpp:55-135). */
int orc_make_grid(const sph_run_config* cfg, orc_grid* g)
{
    memset(g, 0, sizeof(*g));
    if (!(cfg->h0 > 0.0f)) return -1;
    g->support = (cfg->kernel == SPH_KERNEL_QUINTIC) ? 3 : 2;
    g->H = (float)g->support * cfg->h0;
    g->H2 = g->H * g->H;
    g->Hd = (double)g->support * (double)cfg->h0;
    g->H2d = g->Hd * g->Hd;
    const double f = cell_factor_of(cfg->cell_factor);
    g->reach = f < 1.0 ? 2 : 1;
    int64_t ncell = 1;
    for (int a = 0; a < 3; ++a) {
        double L = cfg->box_len[a];
        if (!(L > 0)) return -1;
        double target = (double)g->H * f * (1.0 + 1e-4);
        if (cfg->periodic[a] && L < (double)g->H * (1.0 + 1e-4)) return -1;
        int n = (int)floor(L / target);
        if (n < 1) n = 1;
        g->n[a] = n;
        g->lo[a] = cfg->box_lo[a];
        g->L[a] = L;
        g->s[a] = L / (double)n;
        g->inv_s[a] = (double)n / L;
        g->lo32[a] = (float)g->lo[a];
        g->L32[a] = (float)L;
        g->s32[a] = (float)g->s[a];
        g->inv_s32[a] = (float)g->inv_s[a];
        ncell *= n;
    }
    if (ncell > 0x7fffffffLL) return -1;
    g->ncell = (int32_t)ncell;
    return 0;
}

int orc_bin(const sph_run_config* cfg, int32_t prec, int64_t n, double* const pos[3], uint32_t* key)
{
    orc_grid g;
    if (orc_make_grid(cfg, &g)) return -1;
    for (int64_t i = 0; i < n; ++i) {
        int c[3];
        for (int a = 0; a < 3; ++a) {
            /* clamped before the conversion: outside the box (or not finite) maps to an edge cell */
            int ci;
            if (prec == ORC_PREC_PLAIN) {
                float x = (float)pos[a][i];
                float t = (x - g.lo32[a]) * g.inv_s32[a];
                ci = (int)fminf(fmaxf(floorf(t), 0.0f), (float)(g.n[a] - 1));
            } else {
                double t = (pos[a][i] - g.lo[a]) * g.inv_s[a];
                ci = (int)fmin(fmax(floor(t), 0.0), (double)(g.n[a] - 1));
            }
            c[a] = ci;
        }
        key[i] = (uint32_t)(c[0] + g.n[0] * (c[1] + g.n[1] * c[2]));
    }
    return 0;
}

static const uint64_t* g_sort_ids;
static int cmp_by_id(const void* a, const void* b)
{
    uint64_t x = g_sort_ids[*(const int64_t*)a], y = g_sort_ids[*(const int64_t*)b];
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* Stable order by (supercell key, particle id): a counting sort over keys, then each cell by id. */
int orc_sort(int64_t n, const uint32_t* key, const uint64_t* id, int32_t ncell, int64_t* perm,
             uint32_t* cell_start)
{
    memset(cell_start, 0, ((size_t)ncell + 1) * sizeof(uint32_t));
    for (int64_t i = 0; i < n; ++i) {
        if (key[i] >= (uint32_t)ncell) return -1;
        cell_start[key[i] + 1]++;
    }
    for (int32_t c = 0; c < ncell; ++c) cell_start[c + 1] += cell_start[c];
    uint32_t* cur = (uint32_t*)malloc(((size_t)ncell + 1) * sizeof(uint32_t));
    memcpy(cur, cell_start, ((size_t)ncell + 1) * sizeof(uint32_t));
    for (int64_t i = 0; i < n; ++i) perm[cur[key[i]]++] = i;
    free(cur);
    g_sort_ids = id;
    for (int32_t c = 0; c < ncell; ++c) {
        int64_t b = cell_start[c], e = cell_start[c + 1];
        if (e - b > 1) qsort(perm + b, (size_t)(e - b), sizeof(int64_t), cmp_by_id);
    }
    return 0;
}

/* Region layout of the region traversal: regions are the blocks of block^3 supercells that hold
 * particles at set-up (SetupInterface.hpp:81-96 makes one region per set-up block; here the blocks
 * are supercell-sized so the brute-force search has compact regions to test). A region's particles,
 * in id order, fill its 64-slot frames from region_start[r]; slot[i] is particle i's slot.
 * region_start holds cap_regions + 1 entries. */
int orc_regions_layout(const sph_run_config* cfg, int32_t prec, int64_t n, double* const pos[3], const uint64_t* id,
                       int32_t block, int64_t* slot, int64_t* region_start, int32_t cap_regions, int32_t* n_regions,
                       int64_t* n_slots)
{
    orc_grid g;
    if (orc_make_grid(cfg, &g) || block < 1 || n < 0) return -1;
    int nb[3];
    int64_t nblk = 1;
    for (int a = 0; a < 3; ++a) {
        nb[a] = (g.n[a] + block - 1) / block;
        nblk *= nb[a];
    }
    if (nblk >= 0x7fffffffLL) return -1;
    uint32_t* key = (uint32_t*)malloc((size_t)n * 4 + 4);
    int64_t* perm = (int64_t*)malloc((size_t)n * 8 + 8);
    uint32_t* bs = (uint32_t*)malloc(((size_t)nblk + 1) * 4);
    int rc = (key && perm && bs) ? 0 : -2;
    if (!rc) rc = orc_bin(cfg, prec, n, pos, key);
    if (!rc) {
        for (int64_t i = 0; i < n; ++i) {
            const int cx = (int)(key[i] % (uint32_t)g.n[0]);
            const int cy = (int)((key[i] / (uint32_t)g.n[0]) % (uint32_t)g.n[1]);
            const int cz = (int)(key[i] / ((uint32_t)g.n[0] * (uint32_t)g.n[1]));
            key[i] = (uint32_t)(cx / block + nb[0] * (cy / block + nb[1] * (cz / block)));
        }
        rc = orc_sort(n, key, id, (int32_t)nblk, perm, bs);
    }
    int32_t R = 0;
    int64_t s = 0;
    for (int64_t b = 0; !rc && b < nblk; ++b) {
        const int64_t cnt = (int64_t)bs[b + 1] - (int64_t)bs[b];
        if (cnt == 0) continue;
        if (R >= cap_regions) { rc = -1; break; }
        region_start[R++] = s;
        for (int64_t k = 0; k < cnt; ++k) slot[perm[bs[b] + k]] = s + k;
        s += (cnt + 63) / 64 * 64;
    }
    if (!rc) {
        region_start[R] = s;
        *n_regions = R;
        *n_slots = s;
    }
    free(key);
    free(perm);
    free(bs);
    return rc;
}

#include "sph_oracle_common.h"

#define REAL float
#define S f32
#define RFMA fmaf
#include "sph_oracle_impl.h"
#undef REAL
#undef S
#undef RFMA

#define REAL double
#define S f64
#define RFMA fma
#include "sph_oracle_impl.h"
#undef REAL
#undef S
#undef RFMA
