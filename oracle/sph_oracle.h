/*
 * sph_oracle.h -- CPU oracle for the SPH-Erosion simulation step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / CPU baseline.
 *
 * Two parts:
 *
 *  ref32_*  A faithful FP32 restatement of FluidSystemSPH (the reference's
 *           all-pairs O(N^2) step), array-of-structs, glm operation order,
 *           powf, PI = 3.141592f.  It is pinned bit-for-bit against the
 *           reference's own header compiled into oracle/_ref and against the
 *           golden fixtures under tests/golden/.
 *
 *  orc_*    The library's step as specified for the GPU: canonical (cell, id)
 *           particle order, the reference's FP32 neighbour predicate, the
 *           reference's SPH terms with constants factored out of the sums,
 *           Kahan-compensated or plain sums, symplectic Euler with the
 *           reference's box response, heightfield collision, and the
 *           fixed-point erosion/deposition exchange.  One body is compiled for
 *           double (the accuracy oracle) and for float (a same-order twin of
 *           the GPU kernels).
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ ref32 */

typedef struct { float x, y, z; } orc_v3f;

/* Byte layout of FluidParticle (Erosion/fluid_system.h:49-64): 112 bytes. */
typedef struct {
    int32_t Id;
    orc_v3f Position;
    orc_v3f Velocity;
    orc_v3f Acceleration;
    float Density;
    float Pressure;
    orc_v3f PressureForce;
    orc_v3f ViscosityForce;
    orc_v3f GravityForce;
    orc_v3f SurfaceForce;
    orc_v3f SurfaceNormal;
    int32_t NeighbId;
} ref_particle;

/* Simulation parameters, defaults of Erosion/fluid_system.h:460-478. */
typedef struct {
    float g[3];       /* g = (0,-9.82,0)          :460 */
    float dt;         /* deltaT = 0               :461 */
    float p0;         /* rest density 998.29      :463 */
    float mass;       /* MASS 0.02                :464 */
    float visc;       /* 3.5                      :465 */
    float surf_tens;  /* 0.0728                   :466 */
    float k;          /* stiffness 3              :467 */
    float h;          /* smoothing radius 0.0457  :472 */
    float len;        /* box half extent 0.2      :478 */
} ref_params;

void ref_default_params(ref_params* p);

/* Lattice of Initialize / AddParticles (fluid_system.h:74-102, 232-251).
 * Writes ceil-cbrt(n)^3 records starting with Id = first_id into out (if
 * out != NULL and cap is large enough) and returns the record count. */
int ref_lattice(int n, const float origin[3], int first_id, ref_particle* out, int cap);

/* One FluidSystemSPH::Run (fluid_system.h:104-183) including advance and
 * collisionS (fluid_system.h:306-407).  OpenMP over i; the result does not
 * depend on the thread count. */
void ref_step(ref_particle* P, int n, const ref_params* prm);

/* Timing form of ref_step for the CPU baseline: the three all-pairs passes
 * run only for i = 0, stride, 2*stride, ... (each still over all j) and
 * advance runs for those i only.  Returns the number of i processed. */
int ref_step_sampled(ref_particle* P, int n, const ref_params* prm, int stride);

/* Number of OpenMP threads ref_step uses (1 when built without OpenMP). */
int ref_num_threads(void);

/* ------------------------------------------------------------------ orc */

/* Scene parameters as the host gives them (double; the oracle derives its
 * own float constants from these, independently of the product library). */
typedef struct {
    double h, mass, rho0, k, visc, surf_tens;
    double g[3];
    double dt;
    double box_c[3], box_e;         /* cube collision box: centre, half extent */
    /* terrain (tnx == 0 disables terrain collision and erosion) */
    int32_t tnx, tnz;
    double t_org[3];                /* x0, y0 (terrain base), z0 */
    double t_cell;                  /* terrain cell edge */
    int32_t fx_exp;                 /* fixed-point exponent: 1 unit = 2^-fx_exp volume */
    int32_t erosion;                /* 0 = off, 1 = on */
    double kc, ke, kd, band;        /* capacity coefficient, erode/deposit rates, contact band */
    int32_t lanes;                  /* lanes per particle splitting the neighbour sums (1, 2, 4, 8) */
    int32_t slabs;                  /* x slabs per grid cell (1, 2, 4): order and row clipping */
    int32_t yz;                     /* cells are 1/yz of the x edge along y and z (1, 2 or 3; 0 = 1) */
} orc_scene;

/* Derived per-scene constants, identical in meaning to the GPU's. */
typedef struct {
    float h, h2, thr, t3h2, eps_r, inv_sqrt3;
    float mc6;        /* mass * poly6 constant */
    float pf_scale;   /* mass * spiky constant */
    float vf_scale;   /* visc * viscosity-laplacian constant */
    float ng;         /* -(poly6 gradient constant) */
    float rho0, k, mass, surf_tens, g[3], dt;
    float box_c[3], box_e, box_lo[3], box_hi[3];
    float cell, inv_cell, grid_org[3];
    int32_t dims[3];
    /* terrain */
    float t_org[3], t_inv, hq;  /* height = t_org[1] + (float)vol_fx * hq */
    float kcv, ke, kd, band, up, dn, dmax;
    double dh, dh2, dthr_unused; /* FP64 twins of h, h*h */
    float inv_yz;   /* inv_cell * yz: y and z cell coordinates */
    float yz_scale; /* 1 / yz: y/z cell units to x-cell units */
    int32_t yz;
} orc_consts;

void orc_derive(const orc_scene* s, orc_consts* c);

/* Largest float t with sqrtf(t) <= h: the exact FP32 form of the reference
 * predicate glm::length(r) <= h (fluid_system.h:118-119). */
float orc_pred_threshold(float h);

/* r = |d| and 1/r from d2 = |d|^2 as the FP32 kernels compute them: three
 * Newton steps for 1/sqrt from an integer estimate, r = d2 / sqrt(d2). */
void orc_root_f32(float d2, float* r, float* inv);
void orc_root_f32_n(const float* d2, int n, float* r, float* inv);

/* Half-width of a clipped stencil row before its margin (FP32 kernels). */
float orc_row_width_f32(float x);
void orc_row_width_f32_n(const float* x, int n, float* w);

/* Cell key (x fastest) of each position, FP32. */
void orc_cell_keys(const orc_consts* c, const float* pos3, int n, int32_t* keys);

/* Brute-force neighbour sets under the FP32 predicate (self included):
 * counts[i] and, if idx != NULL, idx[off[i] ...] = ascending j.  Returns the
 * total number of pairs. */
int64_t orc_neighbours_brute(const float* pos3, int n, float thr, int32_t* counts,
                             const int64_t* off, int32_t* idx);

/* Per-particle outputs of a grid step, in input order. */
typedef struct {
    double* rho;        /* density */
    double* acc;        /* 3 per particle */
    double* pforce;     /* pressure force, 3 */
    double* vforce;     /* viscosity force, 3 */
    double* sforce;     /* surface force, 3 */
    double* normal;     /* colour-field normal, 3 */
    int32_t* nbmax;     /* largest neighbour id (j != i), -1 if none */
    int32_t* nbcount;   /* neighbours including self */
} orc_diag;

/* One step of the library's algorithm for one scene, in place, input order
 * preserved.  real = 8 runs the FP64 oracle, real = 4 the FP32 twin.
 * comp (FP32 twin only): 0 plain sums, 1 Kahan-compensated density and force
 * sums, 2 compensated density sums and FP64-accumulated force sums.
 * pos/vel hold 3 doubles per particle (FP32 values when real == 4).
 * vol_fx (tnx*tnz, row = z) and height (float, same shape) are the terrain. */
int orc_step(const orc_scene* s, int real, int comp, int n,
             double* pos, double* vel, int64_t* sed_fx, const int32_t* ids,
             int64_t* vol_fx, float* height, orc_diag* diag);

/* Terrain heights from fixed-point volumes: t_org[1] + (float)vol * hq. */
void orc_heights(const orc_consts* c, int tcells, const int64_t* vol_fx, float* height);
/* Terrain mesh in Grid::UpdateGrid / genIndices form (grid.h:118-176). */
void orc_terrain_mesh(const float* H, int nx, int nz, float cell, const float* org, float* out);
int64_t orc_terrain_indices(int nx, int nz, uint32_t* idx);

/* Erosion / deposition exchange alone (the step's last stage) from a given
 * post-integrate state, in input order.  Returns the number of particles
 * inside the contact band. */
int orc_erode(const orc_scene* s, int n, int real,
              const double* pos, const double* vel, int64_t* sed_fx,
              int64_t* vol_fx, const float* height);

#ifdef __cplusplus
}
#endif
#endif
