/*
 * oracle.c — CPU oracle of the RayTrace per-pixel render path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker, and the timed "port"
 * CPU baseline.  Built with -O3 -ffp-contract=off -fno-fast-math so that every
 * float operation rounds as in the reference's SSE code.
 *
 * Reference files followed: RayTracer.cpp:5-45,450-596,614-696;
 * Basic3DObject.cpp:44-81,135-190,274-300,304-330,376-410,486-552;
 * Model.cpp:402-480,482-664,714-811; 3DElement.cpp:56-238,242-264,408-489.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* Planted mistakes for the pinning tests (built only by the test suite). */
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif
#define MUT(k) (ORACLE_MUTANT == (k))

#define MY_RAY_BASERAY 0x1
#define MY_RAY_SHADOWRAY 0x2
#define MY_RAY_REFLECTRAY 0x3
#define MY_RAY_REFRACTRAY 0x4
#define MY_LIGHT_POINT 0x2

/* Identities compared by the self-skip rules.  Objects use their index; the
 * reference compares object addresses, clTri copy addresses (Model.cpp:775)
 * and `this + cnt` for BallPlane spheres (Basic3DObject.cpp:499). */
#define NO_OBJ ((int64_t)-1)
#define COPY_ID(c) (((int64_t)1 << 40) + (int64_t)(c))
#define TRI_ID(g) (((int64_t)2 << 40) + (int64_t)(g))
#define BALLPLANE_ID(o, cnt) (((int64_t)3 << 40) + (int64_t)(o) * 32 + (cnt))

/* ---------------------------------------------------------------- scene -- */

typedef struct {
    float e1[4], e2[4], p0[4]; /* clTri axisu, axisv, p0 (world) */
    int32_t part;              /* global part index */
    int32_t tri;               /* index within the part (clTri::numb) */
} ocopy;

typedef struct {
    int32_t obj;
    int32_t mtl;
    int32_t first_prim; /* triangle index of the part's first triangle within its model */
    int32_t tri_count;
    int64_t gtri_base;  /* global triangle index of the part's first triangle */
    float bbmin[4], bbmax[4]; /* bboxs (world) */
    int64_t bin_begin[8];
    int32_t bin_count[8];
} opart;

/* global triangle record, REF32 and FP64 forms */
typedef struct {
    float e1[4], e2[4], p0[4];
    double de1[3], de2[3], dp0[3];
    int32_t obj, part, prim;
} otri;

typedef struct {
    double bmin[3], bmax[3];
    int32_t leaf;  /* 1: leaf */
    int32_t a, b;  /* leaf: first, count; inner: children */
} obnode;

typedef struct {
    int32_t kind; /* 0 triangle, 1 sphere */
    int32_t idx;  /* global triangle index / object index */
} oprim;

typedef struct {
    const rtg_scene_desc *d;
    rtg_scene_desc dcopy;
    opart *parts;
    ocopy *copies;
    int64_t n_copies;
    otri *tris;
    int64_t n_tris;
    float (*model_bmin)[4], (*model_bmax)[4]; /* BorderMin/BorderMax per object */
    /* canonical BVH */
    obnode *nodes;
    int64_t n_nodes;
    oprim *prims;
    int64_t n_prims;
} oscene;

struct oracle_scene_s {
    oscene s;
};

typedef struct {
    const oscene *S;
    int max_level;
    int render_mode; /* rtg_params.mode: RTG_MODE_* */
    oracle_stats st;
    oracle_traffic *traffic; /* non-NULL: count canonical-BVH work per traced ray */
    int shadow_matters;      /* traffic only: the pending shadow ray can change the colour */
} octx;

/* ----------------------------------------------------- canonical BVH (FP64) */

typedef struct {
    double bmin[3], bmax[3], c[3];
} pbox;

static void box_empty(double *mn, double *mx) {
    for (int k = 0; k < 3; ++k) mn[k] = 1e300, mx[k] = -1e300;
}
static void box_grow(double *mn, double *mx, const double *a, const double *b) {
    for (int k = 0; k < 3; ++k) {
        if (a[k] < mn[k]) mn[k] = a[k];
        if (b[k] > mx[k]) mx[k] = b[k];
    }
}
static double box_area(const double *mn, const double *mx) {
    double e[3];
    for (int k = 0; k < 3; ++k) e[k] = mx[k] > mn[k] ? mx[k] - mn[k] : 0;
    return 2 * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0]);
}

typedef struct {
    oscene *S;
    pbox *pb;
    int64_t cap;
} bvh_build;

static int64_t bvh_new_node(bvh_build *B) {
    if (B->S->n_nodes == B->cap) {
        B->cap = B->cap ? 2 * B->cap : 1024;
        B->S->nodes = (obnode *)realloc(B->S->nodes, sizeof(obnode) * B->cap);
    }
    return B->S->n_nodes++;
}

/* Binned SAH, 16 bins, leaves of <= 4 primitives. */
static int64_t bvh_rec(bvh_build *B, int64_t first, int64_t count) {
    oscene *S = B->S;
    int64_t ni = bvh_new_node(B);
    double mn[3], mx[3], cmn[3], cmx[3];
    box_empty(mn, mx);
    box_empty(cmn, cmx);
    for (int64_t i = first; i < first + count; ++i) {
        box_grow(mn, mx, B->pb[i].bmin, B->pb[i].bmax);
        box_grow(cmn, cmx, B->pb[i].c, B->pb[i].c);
    }
    obnode *n = &S->nodes[ni];
    memcpy(n->bmin, mn, sizeof(mn));
    memcpy(n->bmax, mx, sizeof(mx));
    if (count <= 4) {
        n->leaf = 1, n->a = (int32_t)first, n->b = (int32_t)count;
        return ni;
    }
    enum { NB = 16 };
    double best_cost = 1e300;
    int best_axis = -1, best_split = -1;
    for (int ax = 0; ax < 3; ++ax) {
        double ext = cmx[ax] - cmn[ax];
        if (!(ext > 0)) continue;
        int64_t cnt[NB] = {0};
        double bmn[NB][3], bmx[NB][3];
        for (int k = 0; k < NB; ++k) box_empty(bmn[k], bmx[k]);
        for (int64_t i = first; i < first + count; ++i) {
            int k = (int)((B->pb[i].c[ax] - cmn[ax]) / ext * NB);
            if (k >= NB) k = NB - 1;
            if (k < 0) k = 0;
            cnt[k]++;
            box_grow(bmn[k], bmx[k], B->pb[i].bmin, B->pb[i].bmax);
        }
        double lmn[3], lmx[3], larea[NB];
        int64_t lcnt[NB];
        box_empty(lmn, lmx);
        int64_t c = 0;
        for (int k = 0; k < NB - 1; ++k) {
            c += cnt[k];
            if (cnt[k]) box_grow(lmn, lmx, bmn[k], bmx[k]);
            larea[k] = c ? box_area(lmn, lmx) : 0;
            lcnt[k] = c;
        }
        double rmn[3], rmx[3];
        box_empty(rmn, rmx);
        c = 0;
        for (int k = NB - 1; k >= 1; --k) {
            c += cnt[k];
            if (cnt[k]) box_grow(rmn, rmx, bmn[k], bmx[k]);
            double cost = larea[k - 1] * (double)lcnt[k - 1] + (c ? box_area(rmn, rmx) * (double)c : 0);
            if (lcnt[k - 1] > 0 && c > 0 && cost < best_cost) best_cost = cost, best_axis = ax, best_split = k;
        }
    }
    int64_t mid;
    if (best_axis < 0) {
        mid = first + count / 2;
    } else {
        double ext = cmx[best_axis] - cmn[best_axis];
        int64_t i = first, j = first + count - 1;
        while (i <= j) {
            int k = (int)((B->pb[i].c[best_axis] - cmn[best_axis]) / ext * NB);
            if (k >= NB) k = NB - 1;
            if (k < best_split)
                ++i;
            else {
                pbox tb = B->pb[i];
                B->pb[i] = B->pb[j];
                B->pb[j] = tb;
                oprim tp = S->prims[i];
                S->prims[i] = S->prims[j];
                S->prims[j] = tp;
                --j;
            }
        }
        mid = i;
        if (mid == first || mid == first + count) mid = first + count / 2;
    }
    int64_t l = bvh_rec(B, first, mid - first);
    int64_t r = bvh_rec(B, mid, first + count - mid);
    n = &S->nodes[ni];
    n->leaf = 0, n->a = (int32_t)l, n->b = (int32_t)r;
    return ni;
}

static void pad_box(pbox *p) {
    double m = 1.0;
    for (int k = 0; k < 3; ++k) {
        if (fabs(p->bmin[k]) > m) m = fabs(p->bmin[k]);
        if (fabs(p->bmax[k]) > m) m = fabs(p->bmax[k]);
    }
    double pad = 1e-5 * m;
    for (int k = 0; k < 3; ++k) p->bmin[k] -= pad, p->bmax[k] += pad, p->c[k] = 0.5 * (p->bmin[k] + p->bmax[k]);
}

static void build_canonical_bvh(oscene *S) {
    const rtg_scene_desc *d = S->d;
    int64_t n = S->n_tris;
    for (int i = 0; i < d->n_objects; ++i)
        if (d->objects[i].show && d->objects[i].type == RTG_OBJ_SPHERE) ++n;
    S->prims = (oprim *)malloc(sizeof(oprim) * (n ? n : 1));
    pbox *pb = (pbox *)malloc(sizeof(pbox) * (n ? n : 1));
    int64_t k = 0;
    for (int64_t t = 0; t < S->n_tris; ++t) {
        const otri *T = &S->tris[t];
        if (!d->objects[T->obj].show) continue;
        double v[3][3];
        for (int c = 0; c < 3; ++c) {
            v[0][c] = T->dp0[c];
            v[1][c] = T->dp0[c] + T->de1[c];
            v[2][c] = T->dp0[c] + T->de2[c];
        }
        box_empty(pb[k].bmin, pb[k].bmax);
        for (int q = 0; q < 3; ++q) box_grow(pb[k].bmin, pb[k].bmax, v[q], v[q]);
        pad_box(&pb[k]);
        S->prims[k].kind = 0, S->prims[k].idx = (int32_t)t;
        ++k;
    }
    for (int i = 0; i < d->n_objects; ++i) {
        const rtg_object *o = &d->objects[i];
        if (!o->show || o->type != RTG_OBJ_SPHERE) continue;
        double c[3] = {o->position.x, o->position.y, o->position.z}, r = fabs((double)o->radius);
        for (int q = 0; q < 3; ++q) pb[k].bmin[q] = c[q] - r, pb[k].bmax[q] = c[q] + r;
        pad_box(&pb[k]);
        S->prims[k].kind = 1, S->prims[k].idx = i;
        ++k;
    }
    S->n_prims = k;
    bvh_build B = {S, pb, 0};
    S->nodes = NULL;
    S->n_nodes = 0;
    if (k > 0) bvh_rec(&B, 0, k);
    free(pb);
}

/* slab test in double; returns 1 on overlap of [0, tmax] */
static inline int dbox(const obnode *n, const double o[3], const double inv[3], double tmax, double *tnear) {
    double t0 = 0, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        double a = (n->bmin[k] - o[k]) * inv[k], b = (n->bmax[k] - o[k]) * inv[k];
        if (a > b) {
            double t = a;
            a = b;
            b = t;
        }
        if (a > t0) t0 = a;
        if (b < t1) t1 = b;
        if (t0 > t1) return 0;
    }
    *tnear = t0;
    return 1;
}

static void ray_inv(const double d[3], double inv[3]) {
    for (int k = 0; k < 3; ++k) {
        double x = d[k];
        if (x == 0) x = signbit(x) ? -1e-300 : 1e-300;
        inv[k] = 1.0 / x;
    }
}

/* Visit every leaf primitive whose box the segment [0, tmax] reaches.  The
 * callback may lower *tmax (closest-hit) or return 1 to stop (any-hit).
 * `visits` counts internal nodes whose child boxes were tested. */
typedef int (*prim_cb)(void *ud, const oprim *p, double *tmax);
static void bvh_walk(const oscene *S, const double o[3], const double d[3], double *tmax, prim_cb cb, void *ud,
                     uint64_t *visits, uint64_t *tri_tests, uint64_t *sph_tests) {
    if (S->n_nodes == 0) return;
    double inv[3];
    ray_inv(d, inv);
    int64_t stack[128];
    int sp = 0;
    double tn;
    if (!dbox(&S->nodes[0], o, inv, *tmax, &tn)) return;
    int64_t cur = 0;
    for (;;) {
        const obnode *n = &S->nodes[cur];
        if (n->leaf) {
            for (int i = 0; i < n->b; ++i) {
                const oprim *p = &S->prims[n->a + i];
                if (p->kind == 0) {
                    if (tri_tests) ++*tri_tests;
                } else if (sph_tests)
                    ++*sph_tests;
                if (cb(ud, p, tmax)) return;
            }
        } else {
            if (visits) ++*visits;
            double ta, tb;
            int ha = dbox(&S->nodes[n->a], o, inv, *tmax, &ta);
            int hb = dbox(&S->nodes[n->b], o, inv, *tmax, &tb);
            if (ha && hb) {
                int64_t nearc = ta <= tb ? n->a : n->b, farc = ta <= tb ? n->b : n->a;
                if (sp < 128) stack[sp++] = farc;
                cur = nearc;
                continue;
            }
            if (ha) {
                cur = n->a;
                continue;
            }
            if (hb) {
                cur = n->b;
                continue;
            }
        }
        if (sp == 0) return;
        cur = stack[--sp];
    }
}

/* ------------------------------------------------------ REF32 instance --- */

#define R float
#define SFX(n) n##32
#define R_SQRT sqrtf
#define R_POW powf
#define R_LOG logf
#define R_EXP expf
#define R_FABS fabsf
#define R_MODF modff
#include "oracle_core.inc"

/* Model::intersect (Model.cpp:748-811) over the octant-binned copies. */
static int model_hit32(const oscene *S, int32_t oi, const RAY32 *ray, const HIT32 *hr, float min, HIT32 *out) {
    const rtg_object *ob = &S->d->objects[oi];
    float empty;
    V32 bmn = v432(S->model_bmin[oi][0], S->model_bmin[oi][1], S->model_bmin[oi][2], S->model_bmin[oi][3]);
    V32 bmx = v432(S->model_bmax[oi][0], S->model_bmax[oi][1], S->model_bmax[oi][2], S->model_bmax[oi][3]);
    float ans = border_test32(ray, bmn, bmx, &empty);
    if (!(ans < hr->distance)) return 0;
    ans = hr->distance;
    int64_t objclt = -1;
    float cw = 0, cu = 0, cv = 0;
    for (int a = 0; a < ob->part_count; ++a) {
        const opart *P = &S->parts[ob->part_begin + a];
        int32_t mask[8];
        V32 pmn = v432(P->bbmin[0], P->bbmin[1], P->bbmin[2], P->bbmin[3]);
        V32 pmx = v432(P->bbmax[0], P->bbmax[1], P->bbmax[2], P->bbmax[3]);
        if (!(border_test_ex32(ray, pmn, pmx, mask) < hr->distance)) continue;
        for (int b = 0; b < 8; ++b) {
            if (!mask[b]) continue;
            for (int64_t c = P->bin_begin[b]; c < P->bin_begin[b] + P->bin_count[b]; ++c) {
                if (hr->obj == COPY_ID(c)) continue;
                const ocopy *T = &S->copies[c];
                float w = 0, u = 0, v = 0;
                float nt = tri_test32(ray, v432(T->e1[0], T->e1[1], T->e1[2], T->e1[3]),
                                      v432(T->e2[0], T->e2[1], T->e2[2], T->e2[3]),
                                      v432(T->p0[0], T->p0[1], T->p0[2], T->p0[3]), &w, &u, &v);
                if (nt < ans) {
                    objclt = c;
                    ans = nt;
                    cw = w, cu = u, cv = v;
                    if (nt < min) goto eos;
                }
            }
        }
    }
eos:
    if (!(ans < hr->distance)) return 0;
    {
        const ocopy *T = &S->copies[objclt];
        const opart *P = &S->parts[T->part];
        const rtg_triangle *tr = &S->d->parts[T->part].tris[T->tri];
        memset(out, 0, sizeof(*out));
        out->distance = ans;
        out->position = vadd32(ray->origin, vscale32(ray->direction, ans));
        V32 n0 = v432(tr->n[0].x, tr->n[0].y, tr->n[0].z, tr->n[0].w);
        V32 n1 = v432(tr->n[1].x, tr->n[1].y, tr->n[1].z, tr->n[1].w);
        V32 n2 = v432(tr->n[2].x, tr->n[2].y, tr->n[2].z, tr->n[2].w);
        out->normal = normalize32(vadd32(vadd32(vscale32(n0, cw), vscale32(n1, cu)), vscale32(n2, cv)));
        out->tu = (tr->uv[0] * cw + tr->uv[2] * cu) + tr->uv[4] * cv;
        out->tv = (tr->uv[1] * cw + tr->uv[3] * cu) + tr->uv[5] * cv;
        out->mtl = P->mtl;
        out->rfr = S->d->materials[P->mtl].rfr;
        out->tex = S->d->materials[P->mtl].tex;
        out->obj = COPY_ID(objclt);
        out->isInside = 0;
        out->id_obj = oi;
        out->id_prim = P->first_prim + T->tri;
    }
    return 1;
}

static int intersect32(const oscene *S, int32_t oi, const RAY32 *ray, const HIT32 *hr, float min, HIT32 *out) {
    const rtg_object *ob = &S->d->objects[oi];
    switch (ob->type) {
    case RTG_OBJ_SPHERE: {
        int r = sphere_hit32(ray, hr, v432(ob->position.x, ob->position.y, ob->position.z, ob->position.w),
                             ob->radius_sqr, ob->radius, oi, &S->d->materials[ob->mtl], ob->mtl, hr->distance, out);
        if (r == 1) out->id_obj = oi, out->id_prim = -1;
        return r != 0;
    }
    case RTG_OBJ_PLANE: return plane_hit32(ray, hr, ob, oi, out);
    case RTG_OBJ_CUBE: return box_hit32(ray, hr, ob, oi, out);
    case RTG_OBJ_BALLPLANE: return ballplane_hit32(ray, hr, ob, oi, &S->d->materials[ob->mtl], out);
    case RTG_OBJ_MODEL: return model_hit32(S, oi, ray, hr, min, out);
    }
    return 0;
}

static void count_ray_traffic(octx *cx, const float o[4], const float d[4], double tmax, int anyhit);

static void closest32(octx *cx, const RAY32 *ray, int64_t base_obj, HIT32 *hr_out, int64_t *newobj_out) {
    const oscene *S = cx->S;
    HIT32 hr;
    memset(&hr, 0, sizeof(hr));
    hr.distance = 1e20f;
    hr.obj = base_obj;
    hr.rfr = 1;
    hr.mtl = -1, hr.tex = -1, hr.id_obj = -1, hr.id_prim = -1;
    int64_t newobj = base_obj;
    for (int i = 0; i < S->d->n_objects; ++i) {
        if (!S->d->objects[i].show) continue;
        hr.obj = base_obj;
        HIT32 h;
        if (intersect32(S, i, ray, &hr, 0, &h)) hr = h;
        if (hr.obj != base_obj) newobj = hr.obj;
    }
    *hr_out = hr;
    *newobj_out = newobj;
    if (cx->traffic) {
        float o[4] = {ray->origin.x, ray->origin.y, ray->origin.z, 0},
              d[4] = {ray->direction.x, ray->direction.y, ray->direction.z, 0};
        count_ray_traffic(cx, o, d, 1e30, 0);
    }
}

static int occluded32(octx *cx, const RAY32 *ray, float dis, int64_t newobj) {
    const oscene *S = cx->S;
    if (cx->traffic) {
        float o[4] = {ray->origin.x, ray->origin.y, ray->origin.z, 0},
              d[4] = {ray->direction.x, ray->direction.y, ray->direction.z, 0};
        count_ray_traffic(cx, o, d, dis, 1);
    }
    HIT32 shr;
    memset(&shr, 0, sizeof(shr));
    shr.distance = dis;
    shr.obj = newobj;
    shr.rfr = 1;
    shr.mtl = -1, shr.tex = -1;
    for (int i = 0; i < S->d->n_objects; ++i) {
        if (S->d->objects[i].show) {
            HIT32 h;
            if (intersect32(S, i, ray, &shr, dis, &h)) shr = h;
        }
        if (shr.distance < dis) return 1;
    }
    return 0;
}

#include "oracle_shade.inc"

#undef R
#undef SFX
#undef R_SQRT
#undef R_POW
#undef R_LOG
#undef R_EXP
#undef R_FABS
#undef R_MODF

/* ------------------------------------------------------- FP64 instance --- */

#define R double
#define SFX(n) n##64
#define R_SQRT sqrt
#define R_POW pow
#define R_LOG log
#define R_EXP exp
#define R_FABS fabs
#define R_MODF modf
#include "oracle_core.inc"

typedef struct {
    const oscene *S;
    const RAY64 *ray;
    int64_t base_obj;
    int anyhit;
    double dis;
    int hit;
    double best_t;
    int64_t best_key; /* tie order: object index, then primitive */
    int32_t best_prim_kind, best_idx;
    double bw, bu, bv;
} walk64;

static int64_t key64(int32_t obj, int32_t prim) { return ((int64_t)obj << 32) + (int64_t)(prim + 1); }

static int prim_cb64(void *ud, const oprim *p, double *tmax) {
    walk64 *W = (walk64 *)ud;
    const oscene *S = W->S;
    double t;
    int32_t obj, prim;
    double w = 0, u = 0, v = 0;
    if (p->kind == 0) {
        const otri *T = &S->tris[p->idx];
        if (W->base_obj == TRI_ID(p->idx)) return 0;
        t = tri_test64(W->ray, v464(T->de1[0], T->de1[1], T->de1[2], 0), v464(T->de2[0], T->de2[1], T->de2[2], 0),
                       v464(T->dp0[0], T->dp0[1], T->dp0[2], 0), &w, &u, &v);
        if (!(t < 1e19)) return 0;
        obj = T->obj, prim = T->prim;
    } else {
        const rtg_object *ob = &S->d->objects[p->idx];
        V64 c = v464(ob->position.x, ob->position.y, ob->position.z, 0);
        double rsq = (double)ob->radius * (double)ob->radius;
        if (W->base_obj == p->idx) {
            if (!W->ray->isInside) return 0;
            t = sphere_exit_t64(W->ray, c, rsq);
        } else
            t = sphere_t64(W->ray, c, rsq);
        if (!(t > 1e-6 && t < 1e19)) return 0;
        obj = p->idx, prim = -1;
    }
    if (W->anyhit) {
        if (t < W->dis) {
            W->hit = 1;
            return 1;
        }
        return 0;
    }
    int64_t k = key64(obj, prim);
    if (t < W->best_t || (t == W->best_t && k < W->best_key)) {
        W->hit = 1;
        W->best_t = t, W->best_key = k;
        W->best_prim_kind = p->kind, W->best_idx = p->idx;
        W->bw = w, W->bu = u, W->bv = v;
        if (t < *tmax) *tmax = t * (1 + 1e-12) + 1e-12;
    }
    return 0;
}

static void closest64(octx *cx, const RAY64 *ray, int64_t base_obj, HIT64 *hr_out, int64_t *newobj_out) {
    const oscene *S = cx->S;
    HIT64 hr;
    memset(&hr, 0, sizeof(hr));
    hr.distance = 1e20f;
    hr.obj = base_obj;
    hr.rfr = 1;
    hr.mtl = -1, hr.tex = -1, hr.id_obj = -1, hr.id_prim = -1;
    int64_t best_key = INT64_MAX;
    /* objects outside the BVH, in scan order */
    for (int i = 0; i < S->d->n_objects; ++i) {
        const rtg_object *ob = &S->d->objects[i];
        if (!ob->show || ob->type == RTG_OBJ_SPHERE || ob->type == RTG_OBJ_MODEL) continue;
        HIT64 probe = hr, h;
        probe.obj = base_obj;
        int r = 0;
        if (ob->type == RTG_OBJ_PLANE) r = plane_hit64(ray, &probe, ob, i, &h);
        if (ob->type == RTG_OBJ_CUBE) r = box_hit64(ray, &probe, ob, i, &h);
        if (ob->type == RTG_OBJ_BALLPLANE) r = ballplane_hit64(ray, &probe, ob, i, &S->d->materials[ob->mtl], &h);
        if (r) hr = h, best_key = key64(i, h.id_prim);
    }
    walk64 W;
    memset(&W, 0, sizeof(W));
    W.S = S, W.ray = ray, W.base_obj = base_obj;
    W.best_t = hr.distance, W.best_key = best_key;
    double o[3] = {ray->origin.x, ray->origin.y, ray->origin.z}, d[3] = {ray->direction.x, ray->direction.y, ray->direction.z};
    double tmax = hr.distance < 1e19 ? hr.distance * (1 + 1e-9) + 1e-9 : 1e30;
    bvh_walk(S, o, d, &tmax, prim_cb64, &W, NULL, NULL, NULL);
    if (W.hit) {
        memset(&hr, 0, sizeof(hr));
        hr.distance = W.best_t;
        hr.position = vadd64(ray->origin, vscale64(ray->direction, W.best_t));
        if (W.best_prim_kind == 0) {
            const otri *T = &S->tris[W.best_idx];
            const rtg_triangle *tr = &S->d->parts[T->part].tris[T->prim - S->parts[T->part].first_prim];
            V64 n0 = v464(tr->n[0].x, tr->n[0].y, tr->n[0].z, 0), n1 = v464(tr->n[1].x, tr->n[1].y, tr->n[1].z, 0),
                n2 = v464(tr->n[2].x, tr->n[2].y, tr->n[2].z, 0);
            hr.normal = normalize64(vadd64(vadd64(vscale64(n0, W.bw), vscale64(n1, W.bu)), vscale64(n2, W.bv)));
            hr.tu = ((double)tr->uv[0] * W.bw + (double)tr->uv[2] * W.bu) + (double)tr->uv[4] * W.bv;
            hr.tv = ((double)tr->uv[1] * W.bw + (double)tr->uv[3] * W.bu) + (double)tr->uv[5] * W.bv;
            const opart *P = &S->parts[T->part];
            hr.mtl = P->mtl;
            hr.rfr = S->d->materials[P->mtl].rfr;
            hr.tex = S->d->materials[P->mtl].tex;
            hr.obj = TRI_ID(W.best_idx);
            hr.isInside = 0;
            hr.id_obj = T->obj, hr.id_prim = T->prim;
        } else {
            const rtg_object *ob = &S->d->objects[W.best_idx];
            V64 c = v464(ob->position.x, ob->position.y, ob->position.z, 0);
            int inside = base_obj == W.best_idx;
            hr.normal = inside ? normalize64(vsub64(c, hr.position)) : normalize64(vsub64(hr.position, c));
            hr.mtl = ob->mtl;
            hr.tex = -1;
            hr.obj = W.best_idx;
            hr.isInside = !ray->isInside;
            hr.rfr = inside && ray->type == MY_RAY_REFRACTRAY ? 1.0 : (double)S->d->materials[ob->mtl].rfr;
            hr.id_obj = W.best_idx, hr.id_prim = -1;
        }
    }
    *hr_out = hr;
    *newobj_out = hr.distance < 1e19 ? hr.obj : base_obj;
}

static int occluded64(octx *cx, const RAY64 *ray, double dis, int64_t newobj) {
    const oscene *S = cx->S;
    for (int i = 0; i < S->d->n_objects; ++i) {
        const rtg_object *ob = &S->d->objects[i];
        if (!ob->show || ob->type == RTG_OBJ_SPHERE || ob->type == RTG_OBJ_MODEL) continue;
        HIT64 probe, h;
        memset(&probe, 0, sizeof(probe));
        probe.distance = dis;
        probe.obj = newobj;
        int r = 0;
        if (ob->type == RTG_OBJ_PLANE) r = plane_hit64(ray, &probe, ob, i, &h);
        if (ob->type == RTG_OBJ_CUBE) r = box_hit64(ray, &probe, ob, i, &h);
        if (ob->type == RTG_OBJ_BALLPLANE) r = ballplane_hit64(ray, &probe, ob, i, &S->d->materials[ob->mtl], &h);
        if (r && h.distance < dis) return 1;
    }
    walk64 W;
    memset(&W, 0, sizeof(W));
    W.S = S, W.ray = ray, W.base_obj = newobj, W.anyhit = 1, W.dis = dis;
    double o[3] = {ray->origin.x, ray->origin.y, ray->origin.z}, d[3] = {ray->direction.x, ray->direction.y, ray->direction.z};
    double tmax = dis;
    bvh_walk(S, o, d, &tmax, prim_cb64, &W, NULL, NULL, NULL);
    return W.hit;
}

#include "oracle_shade.inc"

#undef R
#undef SFX
#undef R_SQRT
#undef R_POW
#undef R_LOG
#undef R_EXP
#undef R_FABS
#undef R_MODF

/* --------------------------------------------------- traffic counting ---- */

typedef struct {
    const oscene *S;
    RAY64 ray;
    int anyhit;
    double dis;
    int hit;
} cwalk;

static int count_cb(void *ud, const oprim *p, double *tmax) {
    cwalk *W = (cwalk *)ud;
    const oscene *S = W->S;
    double t, w, u, v;
    if (p->kind == 0) {
        const otri *T = &S->tris[p->idx];
        t = tri_test64(&W->ray, v464(T->de1[0], T->de1[1], T->de1[2], 0), v464(T->de2[0], T->de2[1], T->de2[2], 0),
                       v464(T->dp0[0], T->dp0[1], T->dp0[2], 0), &w, &u, &v);
    } else {
        const rtg_object *ob = &S->d->objects[p->idx];
        t = sphere_t64(&W->ray, v464(ob->position.x, ob->position.y, ob->position.z, 0),
                       (double)ob->radius * (double)ob->radius);
        if (!(t > 1e-6)) t = 1e20;
    }
    if (!(t < 1e19)) return 0;
    if (W->anyhit) return t < W->dis;
    if (t < *tmax) *tmax = t;
    return 0;
}

static void count_ray_traffic(octx *cx, const float o[4], const float d[4], double tmax, int anyhit) {
    if (anyhit && !cx->shadow_matters) {
        cx->traffic->shadow_skipped++;
        return;
    }
    const uint64_t v0 = cx->traffic->node_visits, t0 = cx->traffic->tri_tests, s0 = cx->traffic->sph_tests;
    cwalk W;
    memset(&W, 0, sizeof(W));
    W.S = cx->S;
    W.ray.origin = v464(o[0], o[1], o[2], 0);
    W.ray.direction = v464(d[0], d[1], d[2], 0);
    W.anyhit = anyhit;
    W.dis = tmax;
    double od[3] = {o[0], o[1], o[2]}, dd[3] = {d[0], d[1], d[2]};
    double tm = tmax;
    cx->traffic->rays++;
    bvh_walk(cx->S, od, dd, &tm, count_cb, &W, &cx->traffic->node_visits, &cx->traffic->tri_tests,
             &cx->traffic->sph_tests);
    if (anyhit) {
        cx->traffic->shadow_rays++;
        cx->traffic->shadow_node_visits += cx->traffic->node_visits - v0;
        cx->traffic->shadow_tri_tests += cx->traffic->tri_tests - t0;
        cx->traffic->shadow_sph_tests += cx->traffic->sph_tests - s0;
    }
}

/* ------------------------------------------------------ near-tie band ---- */

typedef struct {
    const oscene *S;
    RAY32 ray;
    int64_t base_obj;
    int near;
    float t1, t2;
    int64_t k1;
    float win_t;
} ntwalk;

/* FP32 candidates of a primary ray near its winner: flags ties within 4 ulp,
 * barycentric / t margins within 1e-6 of the accept thresholds, and sphere
 * tangency. */
static int nt_cb(void *ud, const oprim *p, double *tmax) {
    ntwalk *W = (ntwalk *)ud;
    const oscene *S = W->S;
    (void)tmax;
    float t;
    int64_t k;
    if (p->kind == 0) {
        const otri *T = &S->tris[p->idx];
        V32 e1 = v432(T->e1[0], T->e1[1], T->e1[2], T->e1[3]), e2 = v432(T->e2[0], T->e2[1], T->e2[2], T->e2[3]);
        V32 p0 = v432(T->p0[0], T->p0[1], T->p0[2], T->p0[3]);
        V32 tmp1 = cross32(W->ray.direction, e2);
        V32 t2r = vsub32(W->ray.origin, p0);
        float f = 1.0f / dot32(e1, tmp1);
        float u = dot32(t2r, tmp1) * f;
        V32 tmp2 = cross32(t2r, e1);
        float v = dot32(W->ray.direction, tmp2) * f, duv = 1.0f - u - v;
        t = dot32(e2, tmp2) * f;
        if (!(t == t) || !(t > 0)) return 0;
        const float eps = 1e-6f;
        int accepted = u >= 0 && u <= 1 && v >= 0 && duv >= 0 && t > 1e-5f;
        int marginal = fabsf(u) <= eps || fabsf(u - 1) <= eps || fabsf(v) <= eps || fabsf(duv) <= eps ||
                       fabsf(t - 1e-5f) <= eps;
        int inside_band = u >= -eps && u <= 1 + eps && v >= -eps && duv >= -eps;
        if (marginal && inside_band && t <= W->win_t * (1 + 1e-5f)) W->near = 1;
        if (!accepted) return 0;
        k = ((int64_t)T->obj << 32) + T->prim + 1;
    } else {
        const rtg_object *ob = &S->d->objects[p->idx];
        V32 c = v432(ob->position.x, ob->position.y, ob->position.z, 0);
        V32 s2r = vsub32(W->ray.origin, c);
        float b = dot32(W->ray.direction, s2r);
        float dis = b * b - dot32(s2r, s2r) + ob->radius_sqr;
        if (fabsf(dis) <= 1e-5f * (ob->radius_sqr + b * b)) W->near = 1;
        t = sphere_t32(&W->ray, c, ob->radius_sqr);
        if (!(t > 1e-6f && t < 1e19f)) return 0;
        k = ((int64_t)p->idx << 32);
    }
    if (t < W->t1 || (t == W->t1 && k < W->k1)) {
        W->t2 = W->t1;
        W->t1 = t, W->k1 = k;
    } else if (t < W->t2)
        W->t2 = t;
    return 0;
}

static uint8_t near_tie_flags(const oscene *S, const RAY32 *ray, const HIT32 *win, float zNear, float zFar) {
    uint8_t fl = 0;
    float wt = win->distance;
    if (wt < 1e19f && (fabsf(wt - zNear) <= 1e-5f * zNear || fabsf(wt - zFar) <= 1e-5f * zFar)) fl |= ORACLE_FLAG_CUT_EDGE;
    ntwalk W;
    memset(&W, 0, sizeof(W));
    W.S = S, W.ray = *ray, W.base_obj = NO_OBJ;
    W.t1 = W.t2 = 1e20f;
    W.k1 = INT64_MAX;
    W.win_t = wt < 1e19f ? wt : 1e30f;
    double o[3] = {ray->origin.x, ray->origin.y, ray->origin.z}, d[3] = {ray->direction.x, ray->direction.y, ray->direction.z};
    double tmax = wt < 1e19f ? (double)wt * (1 + 1e-4) + 1e-4 : 1e30;
    bvh_walk(S, o, d, &tmax, nt_cb, &W, NULL, NULL, NULL);
    /* planes / boxes: parallel band */
    for (int i = 0; i < S->d->n_objects; ++i) {
        const rtg_object *ob = &S->d->objects[i];
        if (!ob->show || ob->type != RTG_OBJ_PLANE) continue;
        V32 N = v432(ob->normal.x, ob->normal.y, ob->normal.z, 0);
        float a = fabsf(dot32(ray->direction, N));
        if (a < 2e-6f && a > 5e-7f) W.near = 1;
    }
    if (W.t1 < 1e19f && W.t2 < 1e19f && W.t2 - W.t1 <= 4 * (nextafterf(W.t1, INFINITY) - W.t1)) W.near = 1;
    /* the binned scan and the brute-force FP32 minimum disagree: culling band */
    int64_t wk = win->id_obj >= 0 ? ((int64_t)win->id_obj << 32) + win->id_prim + 1 : INT64_MAX;
    int win_in_bvh = win->id_obj >= 0 && (S->d->objects[win->id_obj].type == RTG_OBJ_MODEL ||
                                          S->d->objects[win->id_obj].type == RTG_OBJ_SPHERE);
    if (win_in_bvh && (W.t1 != wt || W.k1 != wk)) W.near = 1;
    if (!win_in_bvh && W.t1 < 1e19f && W.t1 <= wt) W.near = 1;
    if (W.near) fl |= ORACLE_FLAG_NEAR_TIE;
    return fl;
}

/* ----------------------------------------------------------- scene prep -- */

int oracle_scene_create(const rtg_scene_desc *d, oracle_scene_t *out) {
    if (d == NULL || out == NULL) return -1;
    struct oracle_scene_s *h = (struct oracle_scene_s *)calloc(1, sizeof(*h));
    oscene *S = &h->s;
    S->dcopy = *d;
    S->d = &S->dcopy;
    S->parts = (opart *)calloc(d->n_parts > 0 ? d->n_parts : 1, sizeof(opart));
    S->model_bmin = calloc(d->n_objects > 0 ? d->n_objects : 1, sizeof(float[4]));
    S->model_bmax = calloc(d->n_objects > 0 ? d->n_objects : 1, sizeof(float[4]));
    S->n_tris = d->n_tris;
    S->tris = (otri *)calloc(d->n_tris > 0 ? d->n_tris : 1, sizeof(otri));
    /* count copies first */
    int64_t n_copies = 0;
    for (int oi = 0; oi < d->n_objects; ++oi) {
        const rtg_object *ob = &d->objects[oi];
        if (ob->type != RTG_OBJ_MODEL) continue;
        for (int a = 0; a < ob->part_count; ++a) n_copies += 8 * (int64_t)d->parts[ob->part_begin + a].tri_count;
    }
    S->copies = (ocopy *)malloc(sizeof(ocopy) * (n_copies > 0 ? n_copies : 1));
    int64_t nc = 0, gt = 0;
    for (int oi = 0; oi < d->n_objects; ++oi) {
        const rtg_object *ob = &d->objects[oi];
        if (ob->type != RTG_OBJ_MODEL) continue;
        /* Model::RTPrepare (Model.cpp:402-480) */
        const float px = ob->position.x, py = ob->position.y, pz = ob->position.z, pw = ob->position.w;
        S->model_bmin[oi][0] = ob->vmin.x + px, S->model_bmin[oi][1] = ob->vmin.y + py;
        S->model_bmin[oi][2] = ob->vmin.z + pz, S->model_bmin[oi][3] = ob->vmin.w + pw;
        S->model_bmax[oi][0] = ob->vmax.x + px, S->model_bmax[oi][1] = ob->vmax.y + py;
        S->model_bmax[oi][2] = ob->vmax.z + pz, S->model_bmax[oi][3] = ob->vmax.w + pw;
        int32_t first_prim = 0;
        for (int a = 0; a < ob->part_count; ++a) {
            const int pi = ob->part_begin + a;
            const rtg_part *rp = &d->parts[pi];
            opart *P = &S->parts[pi];
            P->obj = oi;
            P->mtl = rp->mtl;
            P->first_prim = first_prim;
            P->tri_count = rp->tri_count;
            P->gtri_base = gt;
            P->bbmin[0] = rp->bmin.x + px, P->bbmin[1] = rp->bmin.y + py, P->bbmin[2] = rp->bmin.z + pz,
            P->bbmin[3] = rp->bmin.w + pw;
            P->bbmax[0] = rp->bmax.x + px, P->bbmax[1] = rp->bmax.y + py, P->bbmax[2] = rp->bmax.z + pz,
            P->bbmax[3] = rp->bmax.w + pw;
            /* va = (borders[2p] + borders[2p+1]) * 0.5, model space */
            float mid[3] = {(rp->bmin.x + rp->bmax.x) * 0.5f, (rp->bmin.y + rp->bmax.y) * 0.5f,
                            (rp->bmin.z + rp->bmax.z) * 0.5f};
            int64_t bin_n[8] = {0};
            /* pass 1: sizes */
            for (int t = 0; t < rp->tri_count; ++t) {
                const rtg_triangle *tr = &rp->tris[t];
                float mn[3], mx[3];
                for (int k = 0; k < 3; ++k) {
                    float a0 = (&tr->p[0].x)[k], a1 = (&tr->p[1].x)[k], a2 = (&tr->p[2].x)[k];
                    float m12 = a1 < a2 ? a1 : a2, M12 = a1 > a2 ? a1 : a2;
                    mn[k] = a0 < m12 ? a0 : m12;
                    mx[k] = a0 > M12 ? a0 : M12;
                }
                int in[8] = {0};
                if (mn[0] <= mid[0]) {
                    if (mn[2] <= mid[2]) in[0] = mn[1] <= mid[1], in[1] = mx[1] >= mid[1];
                    if (mx[2] >= mid[2]) in[2] = mn[1] <= mid[1], in[3] = mx[1] >= mid[1];
                }
                if (mx[0] >= mid[0]) {
                    if (mn[2] <= mid[2]) in[4] = mn[1] <= mid[1], in[5] = mx[1] >= mid[1];
                    if (mx[2] >= mid[2]) in[6] = mn[1] <= mid[1], in[7] = mx[1] >= mid[1];
                }
                for (int b = 0; b < 8; ++b) bin_n[b] += in[b];
            }
            for (int b = 0; b < 8; ++b) {
                P->bin_begin[b] = nc;
                P->bin_count[b] = (int32_t)bin_n[b];
                nc += bin_n[b];
            }
            int64_t fill[8];
            for (int b = 0; b < 8; ++b) fill[b] = P->bin_begin[b];
            for (int t = 0; t < rp->tri_count; ++t) {
                const rtg_triangle *tr = &rp->tris[t];
                ocopy C;
                /* clTri(p1 - p0, p2 - p0, p0 + position) (Model.cpp:426) */
                C.e1[0] = tr->p[1].x - tr->p[0].x, C.e1[1] = tr->p[1].y - tr->p[0].y;
                C.e1[2] = tr->p[1].z - tr->p[0].z, C.e1[3] = tr->p[1].w - tr->p[0].w;
                C.e2[0] = tr->p[2].x - tr->p[0].x, C.e2[1] = tr->p[2].y - tr->p[0].y;
                C.e2[2] = tr->p[2].z - tr->p[0].z, C.e2[3] = tr->p[2].w - tr->p[0].w;
                C.p0[0] = tr->p[0].x + px, C.p0[1] = tr->p[0].y + py, C.p0[2] = tr->p[0].z + pz;
                C.p0[3] = tr->p[0].w + pw;
                C.part = pi, C.tri = t;
                otri *T = &S->tris[gt + t];
                memcpy(T->e1, C.e1, sizeof(C.e1));
                memcpy(T->e2, C.e2, sizeof(C.e2));
                memcpy(T->p0, C.p0, sizeof(C.p0));
                for (int k = 0; k < 3; ++k) {
                    T->de1[k] = (double)(&tr->p[1].x)[k] - (double)(&tr->p[0].x)[k];
                    T->de2[k] = (double)(&tr->p[2].x)[k] - (double)(&tr->p[0].x)[k];
                    T->dp0[k] = (double)(&tr->p[0].x)[k] + (double)(&ob->position.x)[k];
                }
                T->obj = oi, T->part = pi, T->prim = first_prim + t;
                float mn[3], mx[3];
                for (int k = 0; k < 3; ++k) {
                    float a0 = (&tr->p[0].x)[k], a1 = (&tr->p[1].x)[k], a2 = (&tr->p[2].x)[k];
                    float m12 = a1 < a2 ? a1 : a2, M12 = a1 > a2 ? a1 : a2;
                    mn[k] = a0 < m12 ? a0 : m12;
                    mx[k] = a0 > M12 ? a0 : M12;
                }
                int in[8] = {0};
                if (mn[0] <= mid[0]) {
                    if (mn[2] <= mid[2]) in[0] = mn[1] <= mid[1], in[1] = mx[1] >= mid[1];
                    if (mx[2] >= mid[2]) in[2] = mn[1] <= mid[1], in[3] = mx[1] >= mid[1];
                }
                if (mx[0] >= mid[0]) {
                    if (mn[2] <= mid[2]) in[4] = mn[1] <= mid[1], in[5] = mx[1] >= mid[1];
                    if (mx[2] >= mid[2]) in[6] = mn[1] <= mid[1], in[7] = mx[1] >= mid[1];
                }
                for (int b = 0; b < 8; ++b)
                    if (in[b]) S->copies[fill[b]++] = C;
            }
            first_prim += rp->tri_count;
            gt += rp->tri_count;
        }
    }
    S->n_copies = nc;
    S->n_tris = gt;
    build_canonical_bvh(S);
    *out = h;
    return 0;
}

void oracle_scene_destroy(oracle_scene_t h) {
    if (h == NULL) return;
    oscene *S = &h->s;
    free(S->parts);
    free(S->copies);
    free(S->tris);
    free(S->model_bmin);
    free(S->model_bmax);
    free(S->nodes);
    free(S->prims);
    free(h);
}

/* ------------------------------------------------------------- render ---- */

static uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint8_t oracle_quantize(float c) {
    /* Color::put (3DElement.cpp:463-468) */
    if (c > 1.0f) return 255;
    if (c < 0.0f) return 0;
    float s = c * 255;
    if (MUT(9)) s = roundf(s);
    return s == s ? (uint8_t)s : 0;
}

typedef struct {
    oscene *S;
    const rtg_params *p;
    int mode, W, H, tiles_x, tiles_y, n_tiles;
    const int32_t *tiles;
    atomic_int next;
    uint8_t *rgb;
    int32_t *ids;
    float *tprim, *rgbf;
    uint8_t *flags;
    double dp;
    float zNear, zFar;
    oracle_traffic *traffic;
    pthread_mutex_t lock;
    oracle_stats total;
} job_t;

static void render_pixel(job_t *J, octx *cx, int x, int y) {
    const rtg_camera *cam = &J->S->d->camera;
    const int xcur = x - J->W / 2, ycur = y - J->H / 2;
    const int spp = J->p->spp > 0 ? J->p->spp : 1;
    if (cx->render_mode == RTG_MODE_CHECK) { /* RTcheck (RayTracer.cpp:46-77): tile checkerboard, no rays */
        const size_t pix = (size_t)y * J->W + x;
        const uint8_t v = ((y / 64) & 1) == ((x / 64) & 1) ? 255 : 0;
        if (J->rgb) J->rgb[3 * pix] = J->rgb[3 * pix + 1] = J->rgb[3 * pix + 2] = v;
        if (J->rgbf) J->rgbf[3 * pix] = J->rgbf[3 * pix + 1] = J->rgbf[3 * pix + 2] = v ? 1.0f : 0.0f;
        if (J->ids) J->ids[2 * pix] = J->ids[2 * pix + 1] = -1;
        if (J->tprim) J->tprim[pix] = 1e20f;
        if (J->flags) J->flags[pix] = 0;
        return;
    }
    float acc32[3] = {0, 0, 0};
    double acc64[3] = {0, 0, 0};
    for (int s = 0; s < spp; ++s) {
        float jx = 0, jy = 0;
        int si = s & 1, sj = (s >> 1) & 1;
        if (spp > 1) {
            uint64_t h = mix64(J->p->seed ^ ((uint64_t)(y * J->W + x) * 0x9E3779B97F4A7C15ull) ^
                               ((uint64_t)s * 0xC2B2AE3D27D4EB4Full));
            jx = (float)(h >> 40) * (1.0f / 16777216.0f);
            jy = (float)((h >> 16) & 0xFFFFFF) * (1.0f / 16777216.0f);
        }
        if (J->mode == ORACLE_REF32) {
            /* parallelRT ray generation (RayTracer.cpp:14,26-27) */
            float a = (float)(xcur * J->dp), b = (float)(ycur * J->dp);
            if (MUT(7)) a = (float)((xcur + 0.5) * J->dp), b = (float)((ycur + 0.5) * J->dp);
            if (spp > 1) {
                float dpf = (float)J->dp;
                a = a + (((float)si + jx) * 0.5f) * dpf;
                b = b + (((float)sj + jy) * 0.5f) * dpf;
            }
            V32 n = v432(cam->n.x, cam->n.y, cam->n.z, cam->n.w), u = v432(cam->u.x, cam->u.y, cam->u.z, cam->u.w),
                v = v432(cam->v.x, cam->v.y, cam->v.z, cam->v.w);
            V32 dir = vadd32(vadd32(n, vscale32(u, a)), vscale32(v, b));
            RAY32 ray;
            ray.origin = v432(cam->position.x, cam->position.y, cam->position.z, cam->position.w);
            ray.direction = normalize32(dir);
            ray.mtlrfr = 1;
            ray.type = MY_RAY_BASERAY;
            ray.isInside = 0;
            HIT32 ph;
            const int dbg = cx->render_mode >= RTG_MODE_DEPTHTEST && cx->render_mode <= RTG_MODE_SHADOWTEST;
            V32 c = dbg ? rtdebug32(cx, cx->render_mode, J->zNear, J->zFar, &ray, &ph)
                        : rtfrac32(cx, J->zNear, J->zFar, &ray, 0, 1.0f, NO_OBJ, &ph);
            acc32[0] += c.x, acc32[1] += c.y, acc32[2] += c.z;
            if (s == 0) {
                size_t pix = (size_t)y * J->W + x;
                if (J->ids) J->ids[2 * pix] = ph.id_obj, J->ids[2 * pix + 1] = ph.id_prim;
                if (J->tprim) J->tprim[pix] = ph.distance;
                if (J->flags) J->flags[pix] = near_tie_flags(J->S, &ray, &ph, J->zNear, J->zFar);
            }
        } else {
            double a = xcur * J->dp, b = ycur * J->dp;
            if (spp > 1) a += (si + (double)jx) * 0.5 * J->dp, b += (sj + (double)jy) * 0.5 * J->dp;
            V64 n = v464(cam->n.x, cam->n.y, cam->n.z, 0), u = v464(cam->u.x, cam->u.y, cam->u.z, 0),
                v = v464(cam->v.x, cam->v.y, cam->v.z, 0);
            V64 dir = vadd64(vadd64(n, vscale64(u, a)), vscale64(v, b));
            RAY64 ray;
            ray.origin = v464(cam->position.x, cam->position.y, cam->position.z, 0);
            ray.direction = normalize64(dir);
            ray.mtlrfr = 1;
            ray.type = MY_RAY_BASERAY;
            ray.isInside = 0;
            HIT64 ph;
            const int dbg = cx->render_mode >= RTG_MODE_DEPTHTEST && cx->render_mode <= RTG_MODE_SHADOWTEST;
            V64 c = dbg ? rtdebug64(cx, cx->render_mode, cam->zNear, 1.4142135623730951 * cam->zFar, &ray, &ph)
                        : rtfrac64(cx, cam->zNear, 1.4142135623730951 * cam->zFar, &ray, 0, 1.0, NO_OBJ, &ph);
            acc64[0] += c.x, acc64[1] += c.y, acc64[2] += c.z;
            if (s == 0) {
                size_t pix = (size_t)y * J->W + x;
                if (J->ids) J->ids[2 * pix] = ph.id_obj, J->ids[2 * pix + 1] = ph.id_prim;
                if (J->tprim) J->tprim[pix] = (float)ph.distance;
                if (J->flags) J->flags[pix] = 0;
            }
        }
    }
    size_t pix = (size_t)y * J->W + x;
    float col[3];
    if (J->mode == ORACLE_REF32) {
        float sc = 1.0f / (float)spp;
        for (int k = 0; k < 3; ++k) col[k] = spp > 1 ? acc32[k] * sc : acc32[k];
    } else {
        for (int k = 0; k < 3; ++k) col[k] = (float)(acc64[k] / spp);
    }
    if (J->rgbf)
        for (int k = 0; k < 3; ++k) J->rgbf[3 * pix + k] = col[k];
    if (J->rgb)
        for (int k = 0; k < 3; ++k) J->rgb[3 * pix + k] = oracle_quantize(col[k]);
}

static void *worker(void *arg) {
    job_t *J = (job_t *)arg;
    octx cx;
    memset(&cx, 0, sizeof(cx));
    cx.S = J->S;
    cx.max_level = J->p->max_level;
    cx.render_mode = J->p->mode;
    oracle_traffic tr;
    memset(&tr, 0, sizeof(tr));
    if (J->traffic) cx.traffic = &tr;
    const int total = J->tiles ? J->n_tiles : J->tiles_x * J->tiles_y;
    for (;;) {
        int k = atomic_fetch_add(&J->next, 1);
        if (k >= total) break;
        int tile = J->tiles ? J->tiles[k] : k;
        int bx = tile % J->tiles_x, by = tile / J->tiles_x;
        if (by >= J->tiles_y) continue;
        int x1 = bx * 64 + 64, y1 = by * 64 + 64;
        if (J->p->tile_compat) {
            if (bx >= J->W / 64 || by >= J->H / 64) continue;
        } else {
            if (x1 > J->W) x1 = J->W;
            if (y1 > J->H) y1 = J->H;
        }
        for (int y = by * 64; y < y1; ++y)
            for (int x = bx * 64; x < x1; ++x) {
                render_pixel(J, &cx, x, y);
                cx.st.pixels++;
            }
    }
    pthread_mutex_lock(&J->lock);
    J->total.rays_primary += cx.st.rays_primary;
    J->total.rays_shadow += cx.st.rays_shadow;
    J->total.rays_reflect += cx.st.rays_reflect;
    J->total.rays_refract += cx.st.rays_refract;
    J->total.pixels += cx.st.pixels;
    if (J->traffic) {
        J->traffic->rays += tr.rays;
        J->traffic->node_visits += tr.node_visits;
        J->traffic->tri_tests += tr.tri_tests;
        J->traffic->sph_tests += tr.sph_tests;
        J->traffic->shadow_rays += tr.shadow_rays;
        J->traffic->shadow_node_visits += tr.shadow_node_visits;
        J->traffic->shadow_tri_tests += tr.shadow_tri_tests;
        J->traffic->shadow_sph_tests += tr.shadow_sph_tests;
        J->traffic->shadow_skipped += tr.shadow_skipped;
    }
    pthread_mutex_unlock(&J->lock);
    return NULL;
}

static int run_job(oracle_scene_t h, const rtg_params *p, int mode, int threads, const int32_t *tiles, int n_tiles,
                   uint8_t *rgb, int32_t *ids, float *tprim, float *rgbf, uint8_t *flags, oracle_traffic *traffic,
                   oracle_stats *stats) {
    if (h == NULL || p == NULL) return -1;
    oscene *S = &h->s;
    job_t J;
    memset(&J, 0, sizeof(J));
    J.S = S, J.p = p, J.mode = mode;
    J.W = p->width > 0 ? p->width : S->d->camera.width;
    J.H = p->height > 0 ? p->height : S->d->camera.height;
    if (J.W <= 0 || J.H <= 0) return -1;
    J.tiles_x = (J.W + 63) / 64, J.tiles_y = (J.H + 63) / 64;
    J.tiles = tiles, J.n_tiles = n_tiles;
    J.rgb = rgb, J.ids = ids, J.tprim = tprim, J.rgbf = rgbf, J.flags = flags;
    J.traffic = traffic;
    /* dp = tan(fovy * PI / 360) / (height / 2), in double (RayTracer.cpp:14) */
    J.dp = tan((double)S->d->camera.fovy * 3.1415926535897932384 / 360) / (J.H / 2);
    J.zNear = S->d->camera.zNear;
    J.zFar = (float)(sqrt(2) * (double)S->d->camera.zFar);
    atomic_init(&J.next, 0);
    pthread_mutex_init(&J.lock, NULL);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    pthread_t th[256];
    for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, worker, &J);
    for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    pthread_mutex_destroy(&J.lock);
    if (stats) {
        *stats = J.total;
        stats->seconds = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
        stats->threads = threads;
    }
    return 0;
}

int oracle_render(oracle_scene_t h, const rtg_params *p, int mode, int threads, const int32_t *tiles, int n_tiles,
                  uint8_t *rgb, int32_t *ids, float *tprim, float *rgbf, uint8_t *flags, oracle_stats *stats) {
    return run_job(h, p, mode, threads, tiles, n_tiles, rgb, ids, tprim, rgbf, flags, NULL, stats);
}

int oracle_traffic_count(oracle_scene_t h, const rtg_params *p, int threads, const int32_t *tiles, int n_tiles,
                         oracle_traffic *out) {
    if (out == NULL || h == NULL) return -1;
    memset(out, 0, sizeof(*out));
    int W = p->width > 0 ? p->width : h->s.d->camera.width, H = p->height > 0 ? p->height : h->s.d->camera.height;
    uint8_t *rgb = (uint8_t *)malloc((size_t)W * H * 3);
    int rc = run_job(h, p, ORACLE_REF32, threads, tiles, n_tiles, rgb, NULL, NULL, NULL, NULL, out, NULL);
    free(rgb);
    out->n_nodes = (uint64_t)h->s.n_nodes;
    out->n_prims = (uint64_t)h->s.n_prims;
    return rc;
}

/* ------------------------------------------------- unit entry points ---- */

static V32 ld4(const float a[4]) { return v432(a[0], a[1], a[2], a[3]); }
float oracle_dot(const float a[4], const float b[4]) { return dot32(ld4(a), ld4(b)); }
void oracle_cross(const float a[4], const float b[4], float out[4]) {
    V32 r = cross32(ld4(a), ld4(b));
    out[0] = r.x, out[1] = r.y, out[2] = r.z, out[3] = r.w;
}
void oracle_normalize(const float a[4], float out[4]) {
    V32 r = normalize32(ld4(a));
    out[0] = r.x, out[1] = r.y, out[2] = r.z, out[3] = r.w;
}
float oracle_tri_test(const float o[4], const float d[4], const float e1[4], const float e2[4], const float p0[4],
                      float bary[3]) {
    RAY32 r;
    r.origin = ld4(o), r.direction = ld4(d);
    float w = 0, u = 0, v = 0;
    float t = tri_test32(&r, ld4(e1), ld4(e2), ld4(p0), &w, &u, &v);
    if (bary) bary[0] = w, bary[1] = u, bary[2] = v;
    return t;
}
float oracle_sphere_test(const float o[4], const float d[4], const float c[4], float radius, float *t_inside) {
    RAY32 r;
    r.origin = ld4(o), r.direction = ld4(d);
    float rsq = radius * radius;
    if (t_inside) *t_inside = sphere_exit_t32(&r, ld4(c), rsq);
    return sphere_t32(&r, ld4(c), rsq);
}
float oracle_plane_test(const float o[4], const float d[4], const float p0[4], const float n[4]) {
    RAY32 r;
    r.origin = ld4(o), r.direction = ld4(d);
    return plane_dist32(&r, ld4(p0), ld4(n));
}
float oracle_border_test(const float o[4], const float d[4], const float bmin[4], const float bmax[4], float *tmax) {
    RAY32 r;
    r.origin = ld4(o), r.direction = ld4(d);
    float tm = 0;
    float t = border_test32(&r, ld4(bmin), ld4(bmax), &tm);
    if (tmax) *tmax = tm;
    return t;
}
float oracle_border_test_ex(const float o[4], const float d[4], const float bmin[4], const float bmax[4],
                            int32_t mask[8]) {
    RAY32 r;
    r.origin = ld4(o), r.direction = ld4(d);
    for (int b = 0; b < 8; ++b) mask[b] = 0;
    return border_test_ex32(&r, ld4(bmin), ld4(bmax), mask);
}
void oracle_raygen(const rtg_camera *cam, int width, int height, int px, int py, float o[4], float d[4]) {
    double dp = tan((double)cam->fovy * 3.1415926535897932384 / 360) / (height / 2);
    int xcur = px - width / 2, ycur = py - height / 2;
    float a = (float)(xcur * dp), b = (float)(ycur * dp);
    if (MUT(7)) a = (float)((xcur + 0.5) * dp), b = (float)((ycur + 0.5) * dp);
    V32 n = v432(cam->n.x, cam->n.y, cam->n.z, cam->n.w), u = v432(cam->u.x, cam->u.y, cam->u.z, cam->u.w),
        v = v432(cam->v.x, cam->v.y, cam->v.z, cam->v.w);
    V32 dir = normalize32(vadd32(vadd32(n, vscale32(u, a)), vscale32(v, b)));
    o[0] = cam->position.x, o[1] = cam->position.y, o[2] = cam->position.z, o[3] = cam->position.w;
    d[0] = dir.x, d[1] = dir.y, d[2] = dir.z, d[3] = dir.w;
}

/* Plane::intersect (Basic3DObject.cpp:376-410) of a ray from another object:
 * the hit distance (1e20: miss) and the texture coordinates. */
float oracle_plane_hit(const float o[4], const float d[4], const float p0[4], const float n[4], const float ax[4],
                       const float ay[4], float uv[2]) {
    RAY32 r;
    memset(&r, 0, sizeof(r));
    r.origin = ld4(o), r.direction = ld4(d), r.type = MY_RAY_BASERAY;
    rtg_object ob;
    memset(&ob, 0, sizeof(ob));
    ob.type = RTG_OBJ_PLANE;
    ob.position = (rtg_vec4){p0[0], p0[1], p0[2], p0[3]};
    ob.normal = (rtg_vec4){n[0], n[1], n[2], n[3]};
    ob.axisx = (rtg_vec4){ax[0], ax[1], ax[2], ax[3]};
    ob.axisy = (rtg_vec4){ay[0], ay[1], ay[2], ay[3]};
    HIT32 hr, out;
    memset(&hr, 0, sizeof(hr));
    hr.obj = NO_OBJ, hr.distance = 1e20f;
    if (!plane_hit32(&r, &hr, &ob, 0, &out)) return 1e20f;
    if (uv) uv[0] = out.tu, uv[1] = out.tv;
    return out.distance;
}

/* Box::intersect (Basic3DObject.cpp:274-300) of a ray from another object:
 * the hit distance (1e20: miss) and the face normal. */
float oracle_box_hit(const float o[4], const float d[4], const float pos[4], const float bmin[4], const float bmax[4],
                     float normal[4]) {
    RAY32 r;
    memset(&r, 0, sizeof(r));
    r.origin = ld4(o), r.direction = ld4(d), r.type = MY_RAY_BASERAY;
    rtg_object ob;
    memset(&ob, 0, sizeof(ob));
    ob.type = RTG_OBJ_CUBE;
    ob.position = (rtg_vec4){pos[0], pos[1], pos[2], pos[3]};
    ob.vmin = (rtg_vec4){bmin[0], bmin[1], bmin[2], bmin[3]};
    ob.vmax = (rtg_vec4){bmax[0], bmax[1], bmax[2], bmax[3]};
    HIT32 hr, out;
    memset(&hr, 0, sizeof(hr));
    hr.obj = NO_OBJ, hr.distance = 1e20f;
    if (!box_hit32(&r, &hr, &ob, 0, &out)) return 1e20f;
    if (normal) normal[0] = out.normal.x, normal[1] = out.normal.y, normal[2] = out.normal.z, normal[3] = out.normal.w;
    return out.distance;
}

/* BallPlane::intersect (Basic3DObject.cpp:486-552) on the frame (pos, ax, ay)
 * with sphere radius r.  The ray leaves sphere `src` (its index in the
 * reference's loop order, cnt - 1; -1: it comes from elsewhere) with the
 * given inside flag and type (MY_RAY_*).  Returns the hit distance (1e20:
 * miss) and, for a hit, out_i = {sphere index, inside flag of the hit};
 * out_f = {normal x, y, z, rfr}. */
float oracle_ballplane_hit(const float o[4], const float d[4], const float pos[4], const float ax[4],
                           const float ay[4], float radius, float mtl_rfr, int src, int inside, int type,
                           int32_t out_i[2], float out_f[4]) {
    RAY32 r;
    memset(&r, 0, sizeof(r));
    r.origin = ld4(o), r.direction = ld4(d), r.type = type, r.isInside = inside ? 1 : 0;
    rtg_object ob;
    memset(&ob, 0, sizeof(ob));
    ob.type = RTG_OBJ_BALLPLANE;
    ob.position = (rtg_vec4){pos[0], pos[1], pos[2], pos[3]};
    ob.axisx = (rtg_vec4){ax[0], ax[1], ax[2], ax[3]};
    ob.axisy = (rtg_vec4){ay[0], ay[1], ay[2], ay[3]};
    ob.radius = radius, ob.radius_sqr = radius * radius;
    rtg_material m;
    memset(&m, 0, sizeof(m));
    m.rfr = mtl_rfr;
    HIT32 hr, out;
    memset(&hr, 0, sizeof(hr));
    hr.obj = src >= 0 ? BALLPLANE_ID(0, src + 1) : NO_OBJ, hr.distance = 1e20f;
    if (!ballplane_hit32(&r, &hr, &ob, 0, &m, &out)) return 1e20f;
    if (out_i) out_i[0] = out.id_prim, out_i[1] = out.isInside;
    if (out_f) out_f[0] = out.normal.x, out_f[1] = out.normal.y, out_f[2] = out.normal.z, out_f[3] = out.rfr;
    return out.distance;
}

/* Color(const Texture*, const Coord2D&) (3DElement.cpp:430-450): the texel
 * colour of texture `tex` (NULL: no texture) at (u, v). */
void oracle_texel(const rtg_texture *tex, float u, float v, float rgb[3]) {
    rtg_scene_desc d;
    memset(&d, 0, sizeof(d));
    d.textures = tex, d.n_textures = tex ? 1 : 0;
    oscene S;
    memset(&S, 0, sizeof(S));
    S.d = &d;
    V32 c = texel32(&S, tex ? 0 : -1, u, v);
    rgb[0] = c.x, rgb[1] = c.y, rgb[2] = c.z;
}
