/*
 * orc_bake.c -- oracle of the per-probe bake loop.  TEST INFRASTRUCTURE
 * ONLY (see vlb_oracle.h).  Compiled with -ffp-contract=off.
 *
 * Per probe (LightBaker::bake, src/baker/light_baker.cpp:287-328, with the
 * coefficient buffer reset per probe):
 *   rays       shaders/env_map.rgen:18-28 (directions from the host table)
 *   hit        shaders/env_map.rchit:51-102 (REFERENCE) or Lambert (LAMBERT)
 *   shadow     env_map.rchit:79-88 + shadow.rmiss:6-9
 *   miss       procedural sky (main.rmiss is unbound in the baker)
 *   project    shaders/sh.comp:23-42 with an exact (double) sum
 */
#include "vlb_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "orc_internal.h"

#define INV_PI_F 0.318309886183790671538f

/* ----------------------------------------------------------------- sky */
void orc_sky(const vlb_sky* sky, const double d[3], double rgb[3])
{
    double g = d[1] > 0.0 ? d[1] : 0.0;
    double sd = d[0] * sky->sun_dir[0] + d[1] * sky->sun_dir[1] + d[2] * sky->sun_dir[2];
    double sun = exp((double)sky->sun_sharpness * (sd - 1.0));
    for (int c = 0; c < 3; ++c)
        rgb[c] = sky->horizon_rgb[c] + (sky->zenith_rgb[c] - sky->horizon_rgb[c]) * g +
                 sky->sun_rgb[c] * sun;
}

static void sky_f(const vlb_sky* sky, const float d[3], float rgb[3])
{
    float g = d[1] > 0.0f ? d[1] : 0.0f;
    float sd = d[0] * sky->sun_dir[0] + d[1] * sky->sun_dir[1] + d[2] * sky->sun_dir[2];
    float sun = expf(sky->sun_sharpness * (sd - 1.0f));
    for (int c = 0; c < 3; ++c)
        rgb[c] = sky->horizon_rgb[c] + (sky->zenith_rgb[c] - sky->horizon_rgb[c]) * g +
                 sky->sun_rgb[c] * sun;
}

/* ------------------------------------------------------------ helpers */
static inline float dot3f(const float a[3], const float b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline double dot3d(const double a[3], const double b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* env_map.rchit:63-68: bc = (1-u-v, u, v), n = sum bc_i n_i, N = normalize(n),
 * hitPosition = o + d t. */
static void hit_geom_f(const orc_scene* s, const float o[3], const float d[3], int prim,
                       float t, float u, float v, float P[3], float N[3])
{
    const float* n = &s->nrm[9 * prim];
    float b0 = 1.0f - u - v;
    float m[3];
    for (int a = 0; a < 3; ++a) {
        P[a] = o[a] + d[a] * t;
        m[a] = n[a] * b0 + n[3 + a] * u + n[6 + a] * v;
    }
    float l = sqrtf(dot3f(m, m));
    for (int a = 0; a < 3; ++a) N[a] = l > 0.0f ? m[a] / l : 0.0f;
}

static void hit_geom_d(const orc_scene* s, const float o[3], const float d[3], int prim,
                       float t, float u, float v, double P[3], double N[3])
{
    const float* n = &s->nrm[9 * prim];
    double b0 = 1.0 - (double)u - (double)v;
    double m[3];
    for (int a = 0; a < 3; ++a) {
        P[a] = (double)o[a] + (double)d[a] * (double)t;
        m[a] = n[a] * b0 + n[3 + a] * (double)u + n[6 + a] * (double)v;
    }
    double l = sqrt(dot3d(m, m));
    for (int a = 0; a < 3; ++a) N[a] = l > 0.0 ? m[a] / l : 0.0;
}

typedef struct {
    const orc_scene* s;
    const vlb_bake_params* p;
    int mode64, brute, K;
    uint32_t R;
    const float* dirs;      /* R*3 */
    const double* wts;      /* R   */
    const double* Yw;       /* R*K or NULL (computed on the fly) */
    double* out;
    float* radiance;
    int32_t* hit;
    volatile uint32_t next;
    uint64_t shadow_total;
    pthread_mutex_t mtx;
    /* multi-bounce: source grid of the indirect term, or ig == NULL */
    const vlb_probe_grid_desc* ig;
    const float* icoeffs;
    const vlb_lookup_params* ilp;
    float iscale;
} bake_job;

/* Radiance of one ray.  rad_f: float pipeline; rad_d: double values with
 * the float pipeline's discrete decisions. */
static void shade_ray(bake_job* J, const float o[3], const float d[3], float rad_f[3],
                      double rad_d[3], int32_t* hit_prim, uint64_t* nshadow)
{
    const orc_scene* s = J->s;
    const vlb_bake_params* p = J->p;
    float t = 0.f, u = 0.f, v = 0.f;
    int prim = s->ntris ? orc_trace_closest(s, o, d, p->t_min, p->t_max, J->brute, &t, &u, &v) : -1;
    *hit_prim = prim;
    const int ref = p->shade_mode == VLB_SHADE_REFERENCE;
    if (prim < 0) {
        float sk[3];
        sky_f(&p->sky, d, sk);
        double dd[3] = {d[0], d[1], d[2]}, skd[3];
        orc_sky(&p->sky, dd, skd);
        for (int c = 0; c < 3; ++c) {
            float x = sk[c];
            double xd = skd[c];
            if (ref && (p->ref.flags & VLB_REF_SRGB)) {
                x = orc_srgb(x);
                xd = xd < 0.0031308 ? xd * 12.92 : 1.055 * pow(xd, 1.0 / 2.4) - 0.055;
            }
            if (ref && (p->ref.flags & VLB_REF_UNORM8)) {
                x = orc_unorm8(x);
                xd = x;
            }
            rad_f[c] = x;
            rad_d[c] = xd;
        }
        return;
    }
    float P[3], N[3];
    double Pd[3], Nd[3];
    hit_geom_f(s, o, d, prim, t, u, v, P, N);
    hit_geom_d(s, o, d, prim, t, u, v, Pd, Nd);
    const float* alb = &s->albedo[3 * prim];
    if (ref) {
        /* env_map.rchit:73-101 / main.rchit:95-116 */
        const vlb_ref_shading* r = &p->ref;
        float L[3], Ln[3];
        for (int a = 0; a < 3; ++a) L[a] = r->light_position[a] - P[a];
        float Llen = sqrtf(dot3f(L, L));
        for (int a = 0; a < 3; ++a) Ln[a] = Llen > 0.0f ? L[a] / Llen : 0.0f;
        float sdn = dot3f(Ln, N);
        float sDotN = sdn > 0.0f ? sdn : 0.0f;
        float org[3];
        for (int a = 0; a < 3; ++a) org[a] = P[a] + r->shadow_bias * N[a];
        int in_shadow = 1;
        if (sDotN != 0.0f) {
            ++*nshadow;
            in_shadow = s->ntris ? orc_trace_occluded(s, org, Ln, 0.0f, Llen, J->brute) : 0;
        }
        float diffuse = 0.0f, spec = 0.0f;
        double diffuse_d = 0.0, spec_d = 0.0;
        if (!in_shadow) {
            diffuse = r->c_diffuse * sDotN;
            float dnl = dot3f(N, Ln);
            float R[3];
            for (int a = 0; a < 3; ++a) R[a] = Ln[a] - 2.0f * dnl * N[a];
            float rd = dot3f(R, d);
            spec = r->c_specular * powf(rd > 0.0f ? rd : 0.0f, r->glossiness);
            /* double values */
            double Ld[3], Lnd[3];
            for (int a = 0; a < 3; ++a) Ld[a] = (double)r->light_position[a] - Pd[a];
            double Lld = sqrt(dot3d(Ld, Ld));
            for (int a = 0; a < 3; ++a) Lnd[a] = Lld > 0.0 ? Ld[a] / Lld : 0.0;
            double sdd = dot3d(Lnd, Nd);
            diffuse_d = r->c_diffuse * (sdd > 0.0 ? sdd : 0.0);
            double dnld = dot3d(Nd, Lnd), Rd[3];
            for (int a = 0; a < 3; ++a) Rd[a] = Lnd[a] - 2.0 * dnld * Nd[a];
            double dd[3] = {d[0], d[1], d[2]};
            double rdd = dot3d(Rd, dd);
            spec_d = r->c_specular * pow(rdd > 0.0 ? rdd : 0.0, r->glossiness);
        }
        int zero = alb[0] == 0.0f && alb[1] == 0.0f && alb[2] == 0.0f;
        for (int c = 0; c < 3; ++c) {
            float base = zero ? 1.0f : alb[c];
            float x = base * (r->ambient + diffuse + spec);
            double xd = (double)base * ((double)r->ambient + diffuse_d + spec_d);
            if (r->flags & VLB_REF_SRGB) {
                x = orc_srgb(x);
                xd = xd < 0.0031308 ? xd * 12.92 : 1.055 * pow(xd, 1.0 / 2.4) - 0.055;
            }
            if (r->flags & VLB_REF_UNORM8) {
                x = orc_unorm8(x);
                xd = x;
            }
            rad_f[c] = x;
            rad_d[c] = xd;
        }
        return;
    }
    /* LAMBERT: two-sided shading normal facing the ray */
    if (dot3f(N, d) > 0.0f) { /* the float decision flips both normals */
        for (int a = 0; a < 3; ++a) {
            N[a] = -N[a];
            Nd[a] = -Nd[a];
        }
    }
    float acc[3] = {0.f, 0.f, 0.f};
    double accd[3] = {0.0, 0.0, 0.0};
    for (uint32_t j = 0; j < p->num_lights && j < VLB_MAX_LIGHTS; ++j) {
        const vlb_light* L = &p->lights[j];
        float Ldir[3], dist, E[3];
        double Ldd[3], Ed[3];
        if (L->type == VLB_LIGHT_POINT) {
            float Lv[3];
            for (int a = 0; a < 3; ++a) Lv[a] = L->vec[a] - P[a];
            dist = sqrtf(dot3f(Lv, Lv));
            for (int a = 0; a < 3; ++a) Ldir[a] = dist > 0.0f ? Lv[a] / dist : 0.0f;
            float d2 = dist * dist;
            for (int c = 0; c < 3; ++c) E[c] = L->rgb[c] / d2;
            double Lvd[3];
            for (int a = 0; a < 3; ++a) Lvd[a] = (double)L->vec[a] - Pd[a];
            double distd = sqrt(dot3d(Lvd, Lvd));
            for (int a = 0; a < 3; ++a) Ldd[a] = distd > 0.0 ? Lvd[a] / distd : 0.0;
            for (int c = 0; c < 3; ++c) Ed[c] = L->rgb[c] / (distd * distd);
        } else {
            for (int a = 0; a < 3; ++a) {
                Ldir[a] = -L->vec[a];
                Ldd[a] = -(double)L->vec[a];
            }
            dist = p->t_max;
            for (int c = 0; c < 3; ++c) {
                E[c] = L->rgb[c];
                Ed[c] = L->rgb[c];
            }
        }
        float cosNL = dot3f(N, Ldir);
        if (cosNL > 0.0f) {
            float org[3];
            for (int a = 0; a < 3; ++a) org[a] = P[a] + p->shadow_bias * N[a];
            ++*nshadow;
            int occ = s->ntris ? orc_trace_occluded(s, org, Ldir, 0.0f, dist, J->brute) : 0;
            if (!occ) {
                double cosd = dot3d(Nd, Ldd);
                for (int c = 0; c < 3; ++c) {
                    acc[c] = acc[c] + E[c] * cosNL;
                    accd[c] += Ed[c] * cosd;
                }
            }
        }
    }
    for (int c = 0; c < 3; ++c) {
        rad_f[c] = alb[c] * acc[c] * INV_PI_F;
        rad_d[c] = (double)alb[c] * accd[c] / ORC_PI_D;
    }
    if (J->ig) {
        /* indirect term: albedo x scale x the grid lookup at the hit point
         * and the two-sided float normal (orc_probe_lookup's arithmetic) */
        double E[3];
        *nshadow += (uint64_t)orc_lookup_point(s, J->ig, J->icoeffs, J->ilp, P, N, E);
        for (int c = 0; c < 3; ++c) {
            const double e = (double)J->iscale * E[c];
            rad_f[c] = rad_f[c] + alb[c] * (float)e;
            rad_d[c] += (double)alb[c] * e;
        }
    }
}

static void bake_probe(bake_job* J, uint32_t pi, uint64_t* nshadow)
{
    const vlb_bake_params* p = J->p;
    const int K = J->K;
    const float* o = &p->probe_xyz[3 * pi];
    double acc[VLB_MAX_K * 3];
    memset(acc, 0, sizeof acc);
    double Yloc[25];
    for (uint32_t r = 0; r < J->R; ++r) {
        const float* d = &J->dirs[3 * r];
        float rf[3];
        double rd[3];
        int32_t hp;
        shade_ray(J, o, d, rf, rd, &hp, nshadow);
        const double* Y;
        if (J->Yw) {
            Y = &J->Yw[(size_t)r * K];
        } else {
            double sh[3] = {d[0], d[2], d[1]};
            orc_sh_eval(K, p->sh_constants == VLB_SH_EXACT, sh, Yloc);
            for (int k = 0; k < K; ++k) Yloc[k] *= J->wts[r];
            Y = Yloc;
        }
        for (int k = 0; k < K; ++k)
            for (int c = 0; c < 3; ++c)
                acc[3 * k + c] += Y[k] * (J->mode64 ? rd[c] : (double)rf[c]);
        size_t ri = (size_t)pi * J->R + r;
        if (J->radiance)
            for (int c = 0; c < 3; ++c) J->radiance[3 * ri + c] = J->mode64 ? (float)rd[c] : rf[c];
        if (J->hit) J->hit[ri] = hp;
    }
    memcpy(&J->out[(size_t)pi * K * 3], acc, sizeof(double) * K * 3);
}

static void* worker(void* arg)
{
    bake_job* J = (bake_job*)arg;
    uint64_t nshadow = 0;
    for (;;) {
        uint32_t pi = __atomic_fetch_add(&J->next, 1u, __ATOMIC_RELAXED);
        if (pi >= J->p->num_probes) break;
        bake_probe(J, pi, &nshadow);
    }
    pthread_mutex_lock(&J->mtx);
    J->shadow_total += nshadow;
    pthread_mutex_unlock(&J->mtx);
    return NULL;
}

static int bake_common(const orc_scene* s, const vlb_bake_params* p, const vlb_probe_grid_desc* ig,
                       const float* icoeffs, const vlb_lookup_params* ilp, float iscale, int mode64,
                       int nthreads, int brute, double* out, float* radiance, int32_t* hit,
                       uint64_t* shadow_rays)
{
    if (!s || !p || !out || !p->probe_xyz) return -1;
    if (p->sh_order != 2 && p->sh_order != 3) return -1;
    const int K = p->sh_order == 2 ? 9 : 16;
    uint32_t R;
    float* dirs;
    double* wts;
    if (p->dir_mode == VLB_DIRS_FIBONACCI) {
        R = p->rays_per_probe;
        if (R == 0) return -1;
        dirs = (float*)malloc((size_t)R * 3 * sizeof(float));
        wts = (double*)malloc((size_t)R * sizeof(double));
        orc_fibonacci(R, dirs);
        for (uint32_t r = 0; r < R; ++r) wts[r] = 4.0 * ORC_PI_D / R;
    } else if (p->dir_mode == VLB_DIRS_EQUIRECT) {
        uint32_t W = p->equirect_w, H = p->equirect_h;
        if (!W || !H) return -1;
        R = W * H;
        dirs = (float*)malloc((size_t)R * 3 * sizeof(float));
        wts = (double*)malloc((size_t)R * sizeof(double));
        float* wf = (float*)malloc((size_t)R * sizeof(float));
        orc_equirect_dirs(W, H, dirs, wf);
        for (uint32_t y = 0; y < H; ++y) {
            double th = ORC_PI_D * (y + 0.5) / H;
            double w = (2.0 * ORC_PI_D / W) * (ORC_PI_D / H) * sin(th);
            for (uint32_t x = 0; x < W; ++x)
                wts[(size_t)y * W + x] = mode64 ? w : (double)wf[(size_t)y * W + x];
        }
        free(wf);
    } else {
        return -1;
    }
    double* Yw = NULL;
    if ((uint64_t)R * K <= (16u << 20)) {
        Yw = (double*)malloc((size_t)R * K * sizeof(double));
        for (uint32_t r = 0; r < R; ++r) {
            const float* d = &dirs[3 * r];
            double sh[3] = {d[0], d[2], d[1]};
            orc_sh_eval(K, p->sh_constants == VLB_SH_EXACT, sh, &Yw[(size_t)r * K]);
            for (int k = 0; k < K; ++k) Yw[(size_t)r * K + k] *= wts[r];
        }
    }
    bake_job J;
    memset(&J, 0, sizeof J);
    J.s = s;
    J.p = p;
    J.mode64 = mode64;
    J.brute = brute;
    J.K = K;
    J.R = R;
    J.dirs = dirs;
    J.wts = wts;
    J.Yw = Yw;
    J.out = out;
    J.radiance = radiance;
    J.hit = hit;
    J.ig = ig;
    J.icoeffs = icoeffs;
    J.ilp = ilp;
    J.iscale = iscale;
    pthread_mutex_init(&J.mtx, NULL);
    if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if (nthreads == 1) {
        worker(&J);
    } else {
        pthread_t th[256];
        for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, worker, &J);
        for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    }
    if (shadow_rays) *shadow_rays = J.shadow_total;
    pthread_mutex_destroy(&J.mtx);
    free(Yw);
    free(dirs);
    free(wts);
    return 0;
}

int orc_bake(const orc_scene* s, const vlb_bake_params* p, int mode64, int nthreads, int brute,
             double* out, float* radiance, int32_t* hit, uint64_t* shadow_rays)
{
    return bake_common(s, p, NULL, NULL, NULL, 0.0f, mode64, nthreads, brute, out, radiance, hit,
                       shadow_rays);
}

int orc_bake_indirect(const orc_scene* s, const vlb_bake_params* p, const vlb_probe_grid_desc* g,
                      const float* coeffs, const vlb_lookup_params* lp, float scale, int mode64,
                      int nthreads, int brute, double* out, float* radiance, int32_t* hit,
                      uint64_t* shadow_rays)
{
    if (!g || !coeffs || !lp || !p || orc_check_grid(g)) return -1;
    if (p->shade_mode != VLB_SHADE_LAMBERT) return -1;
    if (g->num_coefs != (p->sh_order == 2 ? 9u : 16u)) return -1;
    return bake_common(s, p, g, coeffs, lp, scale, mode64, nthreads, brute, out, radiance, hit,
                       shadow_rays);
}
