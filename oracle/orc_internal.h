/* orc_internal.h -- shared internals of the oracle (test infrastructure). */
#ifndef ORC_INTERNAL_H
#define ORC_INTERNAL_H

#include <stdint.h>

#include "vlb_oracle.h"

typedef struct {
    int kx, ky, kz;
    float Sx, Sy, Sz;
} orc_wt;

struct orc_scene {
    uint32_t ntris;
    float* tri;      /* ntris * 9: v0, v1, v2                         */
    float* nrm;      /* ntris * 9: per-corner shading normals          */
    float* albedo;   /* ntris * 3                                       */
    void* nodes;     /* onode array (orc_geom.c), NULL = brute force    */
    uint32_t nnodes;
    uint32_t depth;
    uint32_t* order; /* leaf order -> prim id                           */
    float apad;      /* absolute culling margin                         */
};

void orc_wt_setup(const float d[3], orc_wt* w);
int  orc_wt_hit(const float o[3], const orc_wt* w, const float* v0, const float* v1,
                const float* v2, float tmin, float tmax, float* t, float* u, float* v);
void orc_geometric_normal(const float* v0, const float* v1, const float* v2, float g[3]);

/* orc_lookup.c */
int orc_lookup_point(const orc_scene* s, const vlb_probe_grid_desc* g, const float* coeffs,
                     const vlb_lookup_params* lp, const float p[3], const float nn[3],
                     double out[3]);
int orc_check_grid(const vlb_probe_grid_desc* g);

#define ORC_PI_D 3.14159265358979323846
/* PI of shaders/sh_common.h:1 and main.rmiss:4 */
#define ORC_PI_F 3.1415926538f

#endif
