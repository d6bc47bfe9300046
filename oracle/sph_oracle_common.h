/*
 * sph_oracle_common.h -- helpers shared by the oracle's compiled bodies (TEST INFRASTRUCTURE ONLY):
 * the plain FP32/FP64 build (sph_oracle.c) and the operation-counting build (op_count.cpp).
 */
#ifndef SPH_ORACLE_COMMON_H
#define SPH_ORACLE_COMMON_H

#include "sph_oracle.h"

/* Role predicates (include/spearhed/param/speciesDefinition.param:56-88): targets are Default,
 * sources are Default and Boundary, movable are Default and Tracer. */
static inline int ksupport(int kernel) { return kernel == SPH_KERNEL_QUINTIC ? 3 : 2; }

static inline int is_target(uint8_t sp, uint8_t live) { return live && sp == SPH_SPECIES_DEFAULT; }
static inline int is_source(uint8_t sp, uint8_t live)
{
    return live && (sp == SPH_SPECIES_DEFAULT || sp == SPH_SPECIES_BOUNDARY);
}
static inline int is_movable(uint8_t sp, uint8_t live)
{
    return live && (sp == SPH_SPECIES_DEFAULT || sp == SPH_SPECIES_TRACER);
}

static inline void cell_coords(const orc_grid* g, int cell, int c[3])
{
    c[0] = cell % g->n[0];
    c[1] = (cell / g->n[0]) % g->n[1];
    c[2] = cell / (g->n[0] * g->n[1]);
}

#endif /* SPH_ORACLE_COMMON_H */
