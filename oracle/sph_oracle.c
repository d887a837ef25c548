/*
 * sph_oracle.c -- CPU oracle for the SPH-Erosion simulation step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker and CPU baseline.  The product
 * library never links or calls this file.  Build with -ffp-contract=off and
 * without -ffast-math: the FP32 parts restate exact operation sequences.
 *
 * Parity status per function (what pins it):
 *   ref_lattice        fluid_system.h:74-102, 232-251.  Pinned bit-for-bit
 *                      against oracle/_ref (the reference header compiled
 *                      here) and the committed fixtures lattice_*.
 *   ref_step           fluid_system.h:104-183, 306-407, 410-453.  Pinned
 *                      bit-for-bit against oracle/_ref and fixtures F1-F4,
 *                      plus closed forms (isolated particle, lattice density).
 *   orc_pred_threshold fluid_system.h:118-119 + glm length.  Pinned by an
 *                      exhaustive ulp walk against sqrtf and brute force.
 *   orc_row_width_f32  library-defined row half-width (Newton 1/sqrt, one
 *                      step).  Pinned by a dense sweep of [0, 1] (within
 *                      2e-3 of sqrt) and the clipped-row superset tests.
 *   orc_root_f32       length and normalize of fluid_system.h:118, 438-441
 *                      as the FP32 kernels evaluate them (Newton 1/sqrt).
 *                      Pinned by ulp bounds against sqrtf and 1/sqrtf over
 *                      the FP32 range, and d2 = 0 -> r = 0.
 *   orc_cell_keys /    library-defined grid search.  Pinned by set equality of
 *   orc_step neighbours the grid neighbour sets with orc_neighbours_brute.
 *   orc_step (FP64)    the reference's SPH terms; pinned against ref_step on
 *                      the same states (relative tolerance, different sum
 *                      order and powf) and against closed forms.
 *   orc_step terrain   library-defined heightfield collision and erosion (the
 *   and erosion        reference has none).  Pinned by closed-form single
 *                      particle cases and exact integer mass conservation.
 *   orc_terrain_mesh / grid.h:118-176 (UpdateGrid, genIndices) on a float
 *   orc_terrain_indices heightfield in world units.  Pinned bit-for-bit
 *                      against oracle/_ref (the reference's Grid) and the
 *                      fixture mesh_ref in voxel units (cell 1).
 */
#include "sph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REF_PI 3.141592f /* Erosion/sphere.h:8 (used by every SPH kernel) */

/* ====================================================================== */
/* ref32: faithful restatement of FluidSystemSPH                           */
/* ====================================================================== */

static inline orc_v3f v3(float x, float y, float z) { orc_v3f r = {x, y, z}; return r; }
static inline orc_v3f v3sub(orc_v3f a, orc_v3f b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline orc_v3f v3add(orc_v3f a, orc_v3f b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline orc_v3f v3muls(orc_v3f a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline orc_v3f v3smul(float s, orc_v3f a) { return v3(s * a.x, s * a.y, s * a.z); }
static inline orc_v3f v3neg(orc_v3f a) { return v3(-a.x, -a.y, -a.z); }
static inline orc_v3f v3divs(orc_v3f a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
/* glm dot for vec3: tmp = a*b; tmp.x + tmp.y + tmp.z
 * (vendor/glm/glm/detail/func_geometric.inl:48-55) */
static inline float v3dot(orc_v3f a, orc_v3f b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
/* glm length = sqrt(dot(v, v)) (func_geometric.inl:8-14) */
static inline float v3len(orc_v3f a) { return sqrtf(v3dot(a, a)); }
/* glm normalize = v * inversesqrt(dot(v, v)), inversesqrt = 1 / sqrt
 * (func_geometric.inl:82-90, func_exponential.inl:136-139) */
static inline orc_v3f v3normalize(orc_v3f a) { return v3muls(a, 1.0f / sqrtf(v3dot(a, a))); }

void ref_default_params(ref_params* p)
{
    p->g[0] = 0.0f; p->g[1] = -9.82f; p->g[2] = 0.0f;
    p->dt = 0.0f;
    p->p0 = 998.29f;
    p->mass = 0.02f;
    p->visc = 3.5f;
    p->surf_tens = 0.0728f;
    p->k = 3.0f;
    p->h = 0.0457f;
    p->len = 0.2f;
}

int ref_lattice(int n, const float origin[3], int first_id, ref_particle* out, int cap)
{
    /* fluid_system.h:80-95: i, j, k < cbrt(n); coordinates in double,
     * narrowed to float, then the origin added in float. */
    double cb = cbrt((double)n);
    int side = 0;
    while ((double)side < cb) side++;
    int count = side * side * side;
    if (!out) return count;
    if (cap < count) return -count;
    int id = first_id, w = 0;
    for (int i = 0; i < side; i++)
        for (int j = 0; j < side; j++)
            for (int k = 0; k < side; k++) {
                float x = (float)(-0.2 + i * 0.025);
                float y = (float)(-0.05 + j * 0.025);
                float z = (float)(-0.15 + k * 0.025);
                ref_particle* p = &out[w++];
                memset(p, 0, sizeof(*p));
                p->Id = id++;
                p->Position = v3(x + origin[0], y + origin[1], z + origin[2]);
                p->NeighbId = -1;
            }
    return count;
}

/* kernDefault, fluid_system.h:410-416 */
static inline float ref_kern_default(const ref_params* P, orc_v3f r)
{
    float len = v3len(r);
    if (len > P->h) return 0.0f;
    float c = 315.0f / (64.0f * REF_PI * powf(P->h, 9.0f));
    return c * powf((P->h * P->h - len * len), 3.0f);
}

/* gradDefault, fluid_system.h:418-422 */
static inline orc_v3f ref_grad_default(const ref_params* P, orc_v3f r)
{
    float len = v3len(r);
    float c = 945.0f / (32.0f * REF_PI * powf(P->h, 9.0f));
    return v3neg(v3muls(v3muls(r, c), powf((P->h * P->h - len * len), 2.0f)));
}

/* laplDefault, fluid_system.h:424-428 */
static inline float ref_lapl_default(const ref_params* P, orc_v3f r)
{
    float len = v3len(r);
    float c = 945.0f / (32.0f * REF_PI * powf(P->h, 9.0f));
    float h = P->h;
    return ((-c) * (h * h - len * len)) * (3.0f * h * h - 7.0f * len * len);
}

/* gradPressure, fluid_system.h:430-447 (the r < 10e-5 branch points along
 * normalize(1,1,1), so coincident pairs push both particles the same way). */
static inline orc_v3f ref_grad_pressure(const ref_params* P, orc_v3f r)
{
    float dist = v3len(r);
    float c = 45.f / (REF_PI * powf(P->h, 6.0f));
    orc_v3f dir;
    if ((double)dist < 10e-5)
        dir = v3normalize(v3(1.0f, 1.0f, 1.0f));
    else
        dir = v3normalize(r);
    return v3neg(v3muls(v3muls(dir, c), powf(P->h - dist, 2.0f)));
}

/* laplVisc, fluid_system.h:449-453 */
static inline float ref_lapl_visc(const ref_params* P, orc_v3f r)
{
    float len = v3len(r);
    return (45.0f / (REF_PI * powf(P->h, 6.0f))) * (P->h - len);
}

/* collisionS, fluid_system.h:355-407 (abs is the float overload, as MSVC
 * resolves it). */
static int ref_collision_s(const ref_params* P, orc_v3f pos, orc_v3f* cp, orc_v3f* nrm)
{
    float len = P->len;
    if (fabsf(pos.x) < len && fabsf(pos.y) < len && fabsf(pos.z) < len) return 0;
    char mx = 'x';
    float maxP = fabsf(pos.x);
    if (maxP < fabsf(pos.y)) { mx = 'y'; maxP = fabsf(pos.y); }
    if (maxP < fabsf(pos.z)) { mx = 'z'; maxP = fabsf(pos.z); }
    *cp = pos;
    switch (mx) {
    case 'x':
        if (pos.x < -len) { cp->x = -len; *nrm = v3(1, 0, 0); }
        else              { cp->x = len;  *nrm = v3(-1, 0, 0); }
        break;
    case 'y':
        if (pos.y < -len) { cp->y = -len; *nrm = v3(0, 1, 0); }
        else              { cp->y = len;  *nrm = v3(0, -1, 0); }
        break;
    default:
        if (pos.z < -len) { cp->z = -len; *nrm = v3(0, 0, 1); }
        else              { cp->z = len;  *nrm = v3(0, 0, -1); }
        break;
    }
    return 1;
}

/* Density and pressure of particle i, fluid_system.h:109-124 */
static inline void ref_pass_density(ref_particle* P, int n, const ref_params* prm, int i)
{
    ref_particle* cur = &P[i];
    float density = 0;
    for (int j = 0; j < n; j++) {
        orc_v3f r = v3sub(cur->Position, P[j].Position);
        float distance = v3len(r);
        if (distance <= prm->h) density += prm->mass * ref_kern_default(prm, r);
    }
    cur->Density = density;
    cur->Pressure = prm->k * (cur->Density - prm->p0);
}

/* Pressure, viscosity and colour-field normal, fluid_system.h:129-155 */
static inline void ref_pass_internal(ref_particle* P, int n, const ref_params* prm, int i)
{
    ref_particle* cur = &P[i];
    orc_v3f fPress = v3(0, 0, 0), fVisc = v3(0, 0, 0), nrm = v3(0, 0, 0);
    for (int j = 0; j < n; j++) {
        const ref_particle* nb = &P[j];
        orc_v3f r = v3sub(cur->Position, nb->Position);
        float distance = v3len(r);
        if (distance <= prm->h && i != j) {
            cur->NeighbId = nb->Id;
            float s = (cur->Pressure / (cur->Density * cur->Density) +
                       nb->Pressure / (nb->Density * nb->Density)) * prm->mass;
            fPress = v3add(fPress, v3smul(s, ref_grad_pressure(prm, r)));
            float b = prm->mass / nb->Density;
            fVisc = v3add(fVisc, v3muls(v3muls(v3sub(nb->Velocity, cur->Velocity), b),
                                        ref_lapl_visc(prm, r)));
            nrm = v3add(nrm, v3smul(b, ref_grad_default(prm, r)));
        }
    }
    cur->PressureForce = v3neg(v3muls(fPress, cur->Density));
    cur->ViscosityForce = v3muls(fVisc, prm->visc);
    cur->SurfaceNormal = nrm;
}

/* Gravity and surface tension, fluid_system.h:160-178 */
static inline void ref_pass_external(ref_particle* P, int n, const ref_params* prm, int i)
{
    ref_particle* cur = &P[i];
    cur->GravityForce = v3smul(cur->Density, v3(prm->g[0], prm->g[1], prm->g[2]));
    float cfl = 0.0f;
    for (int j = 0; j < n; j++) {
        orc_v3f r = v3sub(cur->Position, P[j].Position);
        float distance = v3len(r);
        if (distance <= prm->h) cfl += (prm->mass / P[j].Density) * ref_lapl_default(prm, r);
    }
    cur->SurfaceForce = v3smul((-prm->surf_tens) * cfl, cur->SurfaceNormal);
}

/* advance, fluid_system.h:306-353 */
static inline void ref_advance_one(ref_particle* P, const ref_params* prm, int i)
{
    ref_particle* cur = &P[i];
    float dt = prm->dt;
    orc_v3f fInt = v3add(cur->PressureForce, cur->ViscosityForce);
    orc_v3f fExt = v3add(cur->GravityForce, cur->SurfaceForce);
    orc_v3f F = v3add(fInt, fExt);
    orc_v3f acc = v3divs(F, cur->Density);
    cur->Acceleration = acc;
    orc_v3f velNext = v3add(cur->Velocity, v3muls(acc, dt));
    orc_v3f posNext = v3add(cur->Position, v3muls(velNext, dt));
    orc_v3f cp, nrm;
    if (ref_collision_s(prm, posNext, &cp, &nrm) && dt != 0) {
        float d = v3len(v3sub(posNext, cp));
        float s = 1 + 0.5f * d / (dt * v3len(velNext));
        velNext = v3sub(velNext, v3muls(v3muls(nrm, s), v3dot(velNext, nrm)));
        posNext = cp;
    }
    cur->Velocity = velNext;
    cur->Position = posNext;
}

void ref_step(ref_particle* P, int n, const ref_params* prm)
{
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) ref_pass_density(P, n, prm, i);
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) ref_pass_internal(P, n, prm, i);
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) ref_pass_external(P, n, prm, i);
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) ref_advance_one(P, prm, i);
}

int ref_step_sampled(ref_particle* P, int n, const ref_params* prm, int stride)
{
    if (stride < 1) stride = 1;
    int m = (n + stride - 1) / stride, t;
#pragma omp parallel for schedule(static)
    for (t = 0; t < m; t++) ref_pass_density(P, n, prm, t * stride);
#pragma omp parallel for schedule(static)
    for (t = 0; t < m; t++) ref_pass_internal(P, n, prm, t * stride);
#pragma omp parallel for schedule(static)
    for (t = 0; t < m; t++) ref_pass_external(P, n, prm, t * stride);
#pragma omp parallel for schedule(static)
    for (t = 0; t < m; t++) ref_advance_one(P, prm, t * stride);
    return m;
}

int ref_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ====================================================================== */
/* orc: the library's step                                                   */
/* ====================================================================== */

/* r = |d| and 1/r from d2 = |d|^2 as the FP32 kernels compute them
 * (library-defined): three Newton steps y <- y (1.5 - (d2/2) y^2) for 1/sqrt
 * from the integer estimate 0x5f375a86 - bits(d2)/2, each step a multiply,
 * a fused multiply-add and a multiply; r = d2 y.  Only IEEE operations, so
 * it replays the GPU bit for bit.  Pinned against correctly rounded sqrtf
 * and 1/sqrtf over the FP32 range of squared distances (ulp bounds). */
void orc_root_f32(float d2, float* r, float* inv)
{
    int32_t bits;
    memcpy(&bits, &d2, sizeof bits);
    bits = 0x5f375a86 - (bits >> 1);
    float y;
    memcpy(&y, &bits, sizeof y);
    const float hx = 0.5f * d2;
    for (int s = 0; s < 3; s++) {
        const float t = hx * y;
        y = y * fmaf(-t, y, 1.5f);
    }
    *inv = y;
    *r = d2 * y;
}

/* Half-width of a clipped stencil row before its margin, as the FP32 kernels
 * compute it (library-defined): x / sqrt(x) with one Newton step for 1/sqrt(x)
 * from the integer estimate, within 2e-3 of sqrt(x) on [0, 1] (the 2^-6 row
 * margin covers it).  Pinned by a dense sweep of [0, 1] and by the
 * clipped-row superset tests. */
float orc_row_width_f32(float x)
{
    int32_t bits;
    memcpy(&bits, &x, sizeof bits);
    bits = 0x5f375a86 - (bits >> 1);
    float y;
    memcpy(&y, &bits, sizeof y);
    const float t = (0.5f * x) * y;
    y = y * fmaf(-t, y, 1.5f);
    return x * y;
}

void orc_row_width_f32_n(const float* x, int n, float* w)
{
    for (int i = 0; i < n; i++) w[i] = orc_row_width_f32(x[i]);
}

void orc_root_f32_n(const float* d2, int n, float* r, float* inv)
{
    for (int i = 0; i < n; i++) orc_root_f32(d2[i], &r[i], &inv[i]);
}

float orc_pred_threshold(float h)
{
    /* glm::length(r) <= h  <=>  sqrtf(d2) <= h  <=>  d2 <= T, with T the
     * largest float whose correctly rounded square root is <= h (sqrtf is
     * monotone). */
    float t = h * h;
    while (sqrtf(nextafterf(t, INFINITY)) <= h) t = nextafterf(t, INFINITY);
    while (t > 0.0f && sqrtf(t) > h) t = nextafterf(t, 0.0f);
    return t;
}

void orc_derive(const orc_scene* s, orc_consts* c)
{
    memset(c, 0, sizeof(*c));
    const double PI_D = (double)REF_PI;
    float h = (float)s->h;
    double hd = (double)h;
    c->h = h;
    c->h2 = h * h;
    c->thr = orc_pred_threshold(h);
    c->t3h2 = 3.0f * c->h2;
    /* dist < 10e-5 compares the float dist with the double 1e-4
     * (fluid_system.h:438); fl32(1e-4) lies below 1e-4, so the test is
     * dist <= fl32(1e-4), i.e. dist < the next float up. */
    c->eps_r = nextafterf(1e-4f, INFINITY);
    c->inv_sqrt3 = 1.0f / sqrtf(3.0f);
    double mass = (double)(float)s->mass, visc = (double)(float)s->visc;
    double h6 = hd * hd * hd * hd * hd * hd, h9 = h6 * hd * hd * hd;
    c->mc6 = (float)(mass * (315.0 / (64.0 * PI_D * h9)));
    c->pf_scale = (float)(mass * (45.0 / (PI_D * h6)));
    c->vf_scale = (float)(visc * (45.0 / (PI_D * h6)));
    c->ng = (float)(-(945.0 / (32.0 * PI_D * h9)));
    c->rho0 = (float)s->rho0;
    c->k = (float)s->k;
    c->mass = (float)s->mass;
    c->surf_tens = (float)s->surf_tens;
    for (int a = 0; a < 3; a++) c->g[a] = (float)s->g[a];
    c->dt = (float)s->dt;
    c->box_e = (float)s->box_e;
    for (int a = 0; a < 3; a++) {
        c->box_c[a] = (float)s->box_c[a];
        c->box_lo[a] = c->box_c[a] - c->box_e;
        c->box_hi[a] = c->box_c[a] + c->box_e;
    }
    /* cells: h (1 + 2^-8) along x, 1/yz of that along y and z (library-defined) */
    c->yz = s->yz == 2 || s->yz == 3 ? s->yz : 1;
    c->cell = h * 1.00390625f; /* h (1 + 2^-8) */
    c->inv_cell = 1.0f / c->cell;
    c->inv_yz = c->inv_cell * (float)c->yz; /* exact for 1 and 2 */
    c->yz_scale = 1.0f / (float)c->yz;
    for (int a = 0; a < 3; a++) {
        c->grid_org[a] = c->box_lo[a];
        float ext = (c->box_hi[a] - c->box_lo[a]) * (a == 0 ? c->inv_cell : c->inv_yz);
        c->dims[a] = (int)floorf(ext) + 1;
    }
    c->dh = hd;
    c->dh2 = hd * hd;
    if (s->tnx > 0) {
        for (int a = 0; a < 3; a++) c->t_org[a] = (float)s->t_org[a];
        float tc = (float)s->t_cell;
        c->t_inv = 1.0f / tc;
        double area = (double)tc * (double)tc;
        c->hq = (float)(ldexp(1.0, -s->fx_exp) / area);
        c->kcv = (float)((double)(float)s->kc * mass / (double)(float)s->rho0);
        c->ke = (float)s->ke;
        c->kd = (float)s->kd;
        c->band = (float)s->band;
        c->up = (float)ldexp(1.0, s->fx_exp);
        c->dn = (float)ldexp(1.0, -s->fx_exp);
        c->dmax = (float)ldexp(1.0, 52 - s->fx_exp);
    }
}

static inline int clamp_index_f(float f, int nmax)
{
    /* floor already applied; NaN and negatives map to 0 */
    if (!(f >= 0.0f)) return 0;
    if (f >= (float)nmax) return nmax;
    return (int)f;
}

static inline int clamp_index_d(double f, int nmax)
{
    if (!(f >= 0.0)) return 0;
    if (f >= (double)nmax) return nmax;
    return (int)f;
}

void orc_cell_keys(const orc_consts* c, const float* pos3, int n, int32_t* keys)
{
    for (int i = 0; i < n; i++) {
        int k[3];
        for (int a = 0; a < 3; a++) {
            float t = (pos3[3 * i + a] - c->grid_org[a]) * (a == 0 ? c->inv_cell : c->inv_yz);
            k[a] = clamp_index_f(floorf(t), c->dims[a] - 1);
        }
        keys[i] = (k[2] * c->dims[1] + k[1]) * c->dims[0] + k[0];
    }
}

/* Grid::UpdateGrid's vertex stream (grid.h:138-176) for a heightfield of
 * nx x nz cells, row z: (x, y, z, normal) per cell, with x and z at the
 * cell centres org + (i + 0.5) cell and y the height.  The normal is the
 * reference's normalize(cross(u,l) + cross(u,r) + cross(d,l) + cross(d,r))
 * over the vectors to the four edge neighbours present (normalize =
 * v * (1 / sqrt(dot(v, v))), glm func_geometric.inl and
 * func_exponential.inl:136-139); a zero sum gives a zero normal.  With
 * cell 1 and org -0.5 it is the reference's stream on integer heights. */
void orc_terrain_mesh(const float* H, int nx, int nz, float cell, const float* org, float* out)
{
    size_t w = 0;
    for (int z = 0; z < nz; z++)
        for (int x = 0; x < nx; x++) {
            const float y = H[(size_t)z * nx + x];
            float u[3] = {0, 0, 0}, d[3] = {0, 0, 0}, r[3] = {0, 0, 0}, l[3] = {0, 0, 0};
            if (x - 1 >= 0) { l[0] = cell; l[1] = y - H[(size_t)z * nx + x - 1]; }
            if (x + 1 < nx) { r[0] = cell; r[1] = H[(size_t)z * nx + x + 1] - y; }
            if (z - 1 >= 0) { u[1] = y - H[(size_t)(z - 1) * nx + x]; u[2] = cell; }
            if (z + 1 < nz) { d[1] = H[(size_t)(z + 1) * nx + x] - y; d[2] = cell; }
            const float* A[4] = {u, u, d, d};
            const float* B[4] = {l, r, l, r};
            float n[3] = {0, 0, 0};
            for (int k = 0; k < 4; k++) {
                const float* a = A[k];
                const float* b = B[k];
                const float c0 = a[1] * b[2] - a[2] * b[1];
                const float c1 = a[2] * b[0] - a[0] * b[2];
                const float c2 = a[0] * b[1] - a[1] * b[0];
                if (k == 0) { n[0] = c0; n[1] = c1; n[2] = c2; }
                else { n[0] = n[0] + c0; n[1] = n[1] + c1; n[2] = n[2] + c2; }
            }
            const float len = sqrtf((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
            const float inv = len > 0.0f ? 1.0f / len : 0.0f;
            out[w++] = org[0] + ((float)x + 0.5f) * cell;
            out[w++] = y;
            out[w++] = org[2] + ((float)z + 0.5f) * cell;
            out[w++] = n[0] * inv;
            out[w++] = n[1] * inv;
            out[w++] = n[2] * inv;
        }
}

/* Grid::genIndices (grid.h:118-136) for nx x nz vertices; returns the index
 * count (idx may be NULL to count). */
int64_t orc_terrain_indices(int nx, int nz, uint32_t* idx)
{
    int64_t k = 0;
    for (int z = 0, j = nz - 1; z < nz && j >= 0; z++, j--)
        for (int x = 0, i = nx - 1; x < nx && i >= 0; x++, i--) {
            if (x + 1 < nx && z + 1 < nz) {
                if (idx) {
                    idx[k] = (uint32_t)(z * nx + x);
                    idx[k + 1] = (uint32_t)(z * nx + x + 1);
                    idx[k + 2] = (uint32_t)((z + 1) * nx + x);
                }
                k += 3;
            }
            if (j - 1 >= 0 && i - 1 >= 0) {
                if (idx) {
                    idx[k] = (uint32_t)(j * nx + i);
                    idx[k + 1] = (uint32_t)(j * nx + i - 1);
                    idx[k + 2] = (uint32_t)((j - 1) * nx + i);
                }
                k += 3;
            }
        }
    return k;
}

int64_t orc_neighbours_brute(const float* pos3, int n, float thr, int32_t* counts,
                             const int64_t* off, int32_t* idx)
{
    int64_t total = 0;
    int i;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (i = 0; i < n; i++) {
        float xi = pos3[3 * i], yi = pos3[3 * i + 1], zi = pos3[3 * i + 2];
        int cnt = 0;
        for (int j = 0; j < n; j++) {
            float dx = xi - pos3[3 * j], dy = yi - pos3[3 * j + 1], dz = zi - pos3[3 * j + 2];
            float d2 = (dx * dx + dy * dy) + dz * dz;
            if (d2 <= thr) {
                if (idx) idx[off[i] + cnt] = j;
                cnt++;
            }
        }
        counts[i] = cnt;
        total += cnt;
    }
    return total;
}

void orc_heights(const orc_consts* c, int tcells, const int64_t* vol_fx, float* height)
{
    for (int t = 0; t < tcells; t++) {
        float v = (float)vol_fx[t];
        height[t] = c->t_org[1] + v * c->hq;
    }
}

/* ---- canonical order ---- */
typedef struct { int32_t key; int32_t sub; int32_t id; int32_t idx; } orc_sortrec;

/* canonical order: cell key, x slab, id, then input position */
static int cmp_sortrec(const void* a, const void* b)
{
    const orc_sortrec* x = (const orc_sortrec*)a;
    const orc_sortrec* y = (const orc_sortrec*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    if (x->sub != y->sub) return x->sub < y->sub ? -1 : 1;
    if (x->id != y->id) return x->id < y->id ? -1 : 1;
    if (x->idx != y->idx) return x->idx < y->idx ? -1 : 1;
    return 0;
}

/* ---- templated body: double oracle and float twin ---- */
/* candidates per lane per iteration of the GPU density pass */
#define ORC_DENSITY_UNROLL 4
#define OT double
#define OSFX _d
#define OT_ROOT(d2, r, inv) do { (r) = sqrt(d2); (inv) = 1.0 / (r); } while (0)
#define OT_ROW_WIDTH(x) sqrt(x)
#define OT_SQRT sqrt
#define OT_FLOOR floor
#define OT_FABS fabs
#define OT_RINT rint
#include "orc_body.inc"
#undef OT
#undef OSFX
#undef OT_ROOT
#undef OT_ROW_WIDTH
#undef OT_SQRT
#undef OT_FLOOR
#undef OT_FABS
#undef OT_RINT

#define OT float
#define OSFX _f
#define OT_ROOT(d2, r, inv) orc_root_f32((d2), &(r), &(inv))
#define OT_ROW_WIDTH(x) orc_row_width_f32(x)
#define OT_SQRT sqrtf
#define OT_FLOOR floorf
#define OT_FABS fabsf
#define OT_RINT rintf
#include "orc_body.inc"
#undef OT
#undef OSFX
#undef OT_ROOT
#undef OT_ROW_WIDTH
#undef OT_SQRT
#undef OT_FLOOR
#undef OT_FABS
#undef OT_RINT

int orc_step(const orc_scene* s, int real, int comp, int n,
             double* pos, double* vel, int64_t* sed_fx, const int32_t* ids,
             int64_t* vol_fx, float* height, orc_diag* diag)
{
    orc_consts c;
    orc_derive(s, &c);
    if (real == 8) return orc_step_body_d(s, &c, 0, n, pos, vel, sed_fx, ids, vol_fx, height, diag);
    if (real == 4) return orc_step_body_f(s, &c, comp, n, pos, vel, sed_fx, ids, vol_fx, height, diag);
    return -1;
}

int orc_erode(const orc_scene* s, int n, int real,
              const double* pos, const double* vel, int64_t* sed_fx,
              int64_t* vol_fx, const float* height)
{
    orc_consts c;
    orc_derive(s, &c);
    if (s->tnx <= 0 || s->tnz <= 0) return 0;
    if (real == 8) return orc_erode_all_d(s, &c, n, pos, vel, sed_fx, vol_fx, height);
    return orc_erode_all_f(s, &c, n, pos, vel, sed_fx, vol_fx, height);
}
