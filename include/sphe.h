/*
 * sphe.h -- C ABI of the MI355X SPH-Erosion simulation step.
 *
 * Drop-in boundary for the reference's FluidSystemSPH (Erosion/fluid_system.h)
 * and the terrain part of Grid (Erosion/grid.h).  The reference has no FFI:
 * its front end calls the header class directly (Erosion/main.cpp).  Each
 * entry point below names the reference member it replaces; the C++ class in
 * sphe_fluid_system.hpp maps the reference's method names onto them.
 *
 * Conventions: every function returns SPHE_OK (0) or a negative SPHE_E*
 * code; sphe_last_error() gives the message of the last failure on a context.
 * A context is not thread-safe: calls on one context must not overlap.  All
 * device work of a context runs on one HIP stream owned by the context.
 * Host arrays are plain pointers owned by the caller.
 */
#ifndef SPHE_H
#define SPHE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPHE_OK 0
#define SPHE_EARG (-1)    /* invalid argument */
#define SPHE_EHIP (-2)    /* HIP runtime error */
#define SPHE_ESTATE (-3)  /* call not valid in the current state */
#define SPHE_ENOMEM (-4)  /* allocation failed */
#define SPHE_ERANGE (-5)  /* index / id out of range */

/* Arithmetic of the step. */
#define SPHE_FP32 0         /* FP32; Kahan-compensated density and force sums */
#define SPHE_FP32_NAIVE 1   /* FP32, plain sums */
#define SPHE_FP64 2         /* plain FP64 comparison path */
#define SPHE_FP32_F64ACC 3  /* FP32 terms; compensated density sums, force sums accumulated in FP64 */

/* Erosion pass. */
#define SPHE_EROSION_OFF 0
#define SPHE_EROSION_ATOMIC 1    /* 64-bit integer atomics on the heightfield */
#define SPHE_EROSION_TWO_PHASE 2 /* per-particle deltas, then one reduction per terrain cell */

/* Parameter ids for sphe_param_ptr (the reference's pointer getters). */
#define SPHE_P_MASS 0      /* FluidSystemSPH::GetMass    fluid_system.h:261-264 */
#define SPHE_P_VISC 1      /* FluidSystemSPH::GetVisc    fluid_system.h:266-269 */
#define SPHE_P_SURF_TEN 2  /* FluidSystemSPH::GetSurfTen fluid_system.h:271-274 */
#define SPHE_P_P0 3        /* FluidSystemSPH::Getp0      fluid_system.h:276-279 */
#define SPHE_P_GRAV 4      /* FluidSystemSPH::GetGrav    fluid_system.h:281-284 (3 floats) */

/* Per-particle buffer ids for sphe_read_buffer (the reference's FPOS..FPRESS
 * ids, fluid_system.h:22-26, plus the sediment this library adds). */
#define SPHE_BUF_POS 0      /* 3 floats */
#define SPHE_BUF_VEL 1      /* 3 floats */
#define SPHE_BUF_DENSITY 2  /* 1 float (of the last step) */
#define SPHE_BUF_MASS 3     /* 1 float (uniform particle mass) */
#define SPHE_BUF_PRESS 4    /* 1 float (of the last step) */
#define SPHE_BUF_SED 5      /* 1 float, carried sediment volume */
#define SPHE_BUF_ID 6       /* 1 int32 */

/* Simulation parameters of one scene.  sphe_params_default() gives the
 * reference's values (fluid_system.h:460-478). */
typedef struct sphe_params {
    float h;             /* smoothing radius                 (h, 0.0457) */
    float mass;          /* particle mass                    (MASS, 0.02) */
    float rest_density;  /* rest density                     (p0, 998.29) */
    float stiffness;     /* pressure constant                (k, 3) */
    float visc;          /* viscosity coefficient            (visc, 3.5) */
    float surf_tens;     /* surface tension coefficient      (surf_tens, 0.0728) */
    float gravity[3];    /* gravity                          (g, (0,-9.82,0)) */
    float dt;            /* time step                        (deltaT, 0) */
    float box_center[3]; /* collision cube centre            (0,0,0) */
    float box_half;      /* collision cube half extent       (len, 0.2) */
    /* erosion model (this library's; the reference has none) */
    int32_t erosion;     /* SPHE_EROSION_* */
    float kc;            /* sediment capacity per (m/s) of speed, times particle volume m/rho0 */
    float ke;            /* erode rate per step */
    float kd;            /* deposit rate per step */
    float band;          /* contact band above the terrain surface */
} sphe_params;

/* Heightfield of one scene: nx * nz cells of edge `cell`, cell (i, k)
 * centred at (origin[0] + (i+0.5) cell, origin[2] + (k+0.5) cell), height
 * heights[k * nx + i] (absolute y, >= origin[1]).  nx == 0: no terrain. */
typedef struct sphe_terrain_desc {
    int32_t nx, nz;
    float origin[3];
    float cell;
    const float* heights;
} sphe_terrain_desc;

/* FluidParticle (fluid_system.h:49-64), 112 bytes, as GetParticle returns it. */
typedef struct sphe_particle {
    int32_t Id;
    float Position[3];
    float Velocity[3];
    float Acceleration[3];
    float Density;
    float Pressure;
    float PressureForce[3];
    float ViscosityForce[3];
    float GravityForce[3];
    float SurfaceForce[3];
    float SurfaceNormal[3];
    int32_t NeighbId;
} sphe_particle;

typedef struct sphe_ctx sphe_ctx;

/* ---- lifetime ---------------------------------------------------------- */

/* FluidSystemSPH::FluidSystemSPH (fluid_system.h:69-72).  precision is one of
 * SPHE_FP32, SPHE_FP32_NAIVE, SPHE_FP64, SPHE_FP32_F64ACC.  Uses the current
 * HIP device. */
int sphe_create(int precision, sphe_ctx** out);
int sphe_destroy(sphe_ctx* ctx);
const char* sphe_last_error(const sphe_ctx* ctx);
/* Library version string. */
const char* sphe_version(void);

void sphe_params_default(sphe_params* p);

/* Adds a scene (one FluidSystemSPH + its terrain).  Scenes advance together
 * in each sphe_step.  terrain may be NULL. */
int sphe_add_scene(sphe_ctx* ctx, const sphe_params* p, const sphe_terrain_desc* terrain, int* scene);
int sphe_scene_count(const sphe_ctx* ctx, int* count);

/* ---- parameters -------------------------------------------------------- */

int sphe_set_params(sphe_ctx* ctx, int scene, const sphe_params* p);
int sphe_get_params(const sphe_ctx* ctx, int scene, sphe_params* p);
/* Stable host storage of one parameter (SPHE_P_*), read at every step, so a
 * UI may write through it between steps like the reference's getters. */
float* sphe_param_ptr(sphe_ctx* ctx, int scene, int which);
/* FluidSystemSPH::SetDeltaTime / GetDeltaTime (fluid_system.h:216-224) */
int sphe_set_dt(sphe_ctx* ctx, int scene, float dt);
int sphe_get_dt(const sphe_ctx* ctx, int scene, float* dt);
/* FluidSystemSPH::SetOrigin / GetOrigin (fluid_system.h:206-214) */
int sphe_set_origin(sphe_ctx* ctx, int scene, const float origin[3]);
int sphe_get_origin(const sphe_ctx* ctx, int scene, float origin[3]);

/* ---- particles --------------------------------------------------------- */

/* FluidSystemSPH::Initialize (fluid_system.h:74-102): appends the reference
 * lattice of ceil(cbrt(n))^3 particles at the scene origin. */
int sphe_init_lattice(sphe_ctx* ctx, int scene, int n);
/* FluidSystemSPH::AddParticles (fluid_system.h:232-251) */
int sphe_add_particles(sphe_ctx* ctx, int scene, int n);
/* FluidSystemSPH::Reset (fluid_system.h:253-259) */
int sphe_reset(sphe_ctx* ctx, int scene);
/* Replaces a scene's particles (any state; used for synthetic scenes and
 * checkpoints).  vel, ids, sed may be NULL (zero velocity, ids 0..n-1, zero
 * sediment).  ids must be unique and non-negative. */
int sphe_set_particles(sphe_ctx* ctx, int scene, int n, const float* pos3, const float* vel3,
                       const int32_t* ids, const int64_t* sed_fx);
/* Same in FP64 (for the FP64 context). */
int sphe_set_particles_f64(sphe_ctx* ctx, int scene, int n, const double* pos3, const double* vel3,
                           const int32_t* ids, const int64_t* sed_fx);
int sphe_num_particles(const sphe_ctx* ctx, int scene, int* n);
/* Copies one per-particle buffer (SPHE_BUF_*) in ascending id order.  After
 * a step, positions, velocities, densities and pressures of a scene whose
 * ids are 0 .. n-1 are gathered in id order on the device and copied alone;
 * positions, velocities and sediment of other scenes bring the whole
 * context's state to the host once per step. */
int sphe_read_buffer(sphe_ctx* ctx, int scene, int which, void* out, size_t bytes);
/* The same buffer as a host view owned by the context: *view points at
 * *count records (3 floats for POS and VEL, else 1 float or int32), a
 * snapshot that stays valid until the next sphe_buffer call for the same
 * scene and buffer or sphe_destroy.  The reference hands out views of its
 * per-particle fields (FluidParticle members through GetParticle,
 * fluid_system.h:286-289; the FPOS .. FPRESS field ids, :22-26). */
int sphe_buffer(sphe_ctx* ctx, int scene, int which, const void** view, size_t* count);
/* Positions and velocities in FP64 (FP64 context) in ascending id order. */
int sphe_read_state_f64(sphe_ctx* ctx, int scene, double* pos3, double* vel3);
/* Fixed-point sediment (1 unit = 2^-sphe_fixed_exponent volume), id order. */
int sphe_read_sediment_fx(sphe_ctx* ctx, int scene, int64_t* sed_fx);
/* FluidSystemSPH::GetParticle (fluid_system.h:286-289): the 112-byte record
 * of particle `id` after the last step (forces of that step, summed in the
 * order the step summed them).  NeighbId is the largest neighbour id, -1 when
 * the particle had no neighbour.  Costs one small kernel and one 232-byte
 * read-back; the first call after a step also rebuilds the context's
 * id -> slot map on the device (one pass over the slots).  Ids that are
 * sparse (largest id above 2 n + 4096) are found by a search over the
 * scene's slots instead.  With duplicate ids (Initialize on a populated
 * scene) the record is one of them.  Before any step (or after an edit) the
 * record has the position and velocity and zero forces. */
int sphe_get_particle(sphe_ctx* ctx, int scene, int id, sphe_particle* out);

/* ---- terrain ----------------------------------------------------------- */

/* Attaches (or replaces, or with nx == 0 removes) a scene's heightfield;
 * heights may be NULL for a flat terrain at origin[1].  The reference's
 * front end holds its heightfield in Grid (grid.h:26-51) and never couples it
 * to the fluid; this is the coupling. */
int sphe_set_terrain(sphe_ctx* ctx, int scene, const sphe_terrain_desc* terrain);
/* Grid::LoadHeightfield (grid.h:98-102): uint8 heights of a row-major
 * image of w pixels per row and h rows (stbi_load's layout, main.cpp:88-92).
 * Terrain cell (x, z) takes Grid::GetHeightfieldAt(x, z) = img[w * x + z]
 * (grid.h:104-107: x picks the row, z the column), so the terrain has h cells
 * along x and w along z; height = origin y + scale * pixel.  The terrain's
 * origin and cell edge are kept (cell 1 if none was set). */
int sphe_terrain_load_u8(sphe_ctx* ctx, int scene, const uint8_t* img, int w, int h, float scale);
/* Heights (absolute y) of the scene's cells, row = z. */
int sphe_terrain_heights(sphe_ctx* ctx, int scene, float* out, size_t count);
/* Fixed-point cell volumes, row = z. */
int sphe_terrain_volume_fx(sphe_ctx* ctx, int scene, int64_t* out, size_t count);
/* Restores fixed-point cell volumes read by sphe_terrain_volume_fx (count =
 * nx * nz, |v| <= 2^62, the scene's fixed exponent unchanged; erosion is
 * not capped at a cell's volume, so a cell may be below the terrain base): with
 * sphe_set_particles_f64 (sediment included) a checkpoint of a scene, from
 * which the steps continue bit for bit.  No counterpart in the reference,
 * whose terrain is static. */
int sphe_set_terrain_volume_fx(sphe_ctx* ctx, int scene, const int64_t* vol, size_t count);
int sphe_terrain_dims(const sphe_ctx* ctx, int scene, int* nx, int* nz);
/* Exponent e of the scene's fixed-point volume unit 2^-e. */
int sphe_fixed_exponent(const sphe_ctx* ctx, int scene, int* e);
/* Grid::UpdateGrid vertex stream (grid.h:138-176): 6 floats per cell
 * (x, y, z, nx, ny, nz) of the eroded terrain, and the triangle indices of
 * Grid::genIndices (grid.h:118-136).  Pass NULL to query sizes. */
int sphe_terrain_mesh(sphe_ctx* ctx, int scene, float* verts, size_t* nverts_floats,
                      uint32_t* indices, size_t* nindices);
/* The same vertex stream left in device memory (computed by a kernel from
 * the current heights) for a renderer that copies it into its vertex buffer
 * through graphics interop: *dverts stays valid until the next call that
 * steps, edits or destroys the context, or asks for another mesh.  Replaces
 * the host round trip of Grid::UpdateGrid + glBufferData
 * (grid.h:138-176, main.cpp:228-240). */
int sphe_terrain_mesh_device(sphe_ctx* ctx, int scene, const float** dverts, size_t* nfloats);
/* The scene's particle positions, 3 floats each in ascending id order, left
 * in device memory for a renderer that copies them into an instance buffer
 * through graphics interop and draws every particle in one instanced call.
 * Replaces FluidSystemSPH::Draw's one draw call per particle
 * (fluid_system.h:185-204).  The scene's ids must be 0 .. n-1 (Initialize,
 * AddParticles and Reset give them; SPHE_EARG otherwise).  *dpos stays valid
 * until the next call that steps, edits or destroys the context, or reads a
 * gathered buffer again. */
int sphe_positions_device(sphe_ctx* ctx, int scene, const float** dpos, size_t* nfloats);

/* ---- stepping ---------------------------------------------------------- */

/* FluidSystemSPH::Run (fluid_system.h:104-183) for every scene, nsteps
 * times.  Returns after the work is queued; sphe_sync waits for it. */
int sphe_step(sphe_ctx* ctx, int nsteps);
/* A batch: one context holding n scenes (params[i], terrains[i] or none if
 * terrains is NULL), which every step advances together, one launch per
 * kernel for all of them.  On an error nothing is created.  No counterpart
 * in the reference (one FluidSystemSPH per scene). */
int sphe_batch_create(int precision, const sphe_params* params, const sphe_terrain_desc* terrains, int n,
                      sphe_ctx** out);
/* Advances every scene of a batch nsteps times: sphe_step. */
int sphe_batch_step(sphe_ctx* ctx, int nsteps);
int sphe_sync(sphe_ctx* ctx);
/* Capture the step sequence in a HIP graph (default on). */
int sphe_use_graph(sphe_ctx* ctx, int enable);

/* Water and sediment bookkeeping of one scene, reduced on the GPU in FP64:
 * terrain volume, carried sediment, and the exact mod-2^64 sum of all
 * fixed-point volume units (conserved exactly by every step). */
int sphe_mass(sphe_ctx* ctx, int scene, double* terrain, double* sediment, uint64_t* units);
/* The same for every scene of the context in one launch (per scene a fixed
 * number of blocks whose partials are added in block order, so the sums are
 * deterministic): terrain[s], sediment[s], units[s] for each scene s; the
 * arrays hold count >= the scene count entries (NULL: not wanted).  No
 * counterpart in the reference. */
int sphe_mass_all(sphe_ctx* ctx, double* terrain, double* sediment, uint64_t* units, int count);
/* sphe_mass_all without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Its partial sums and
 * results live in a static scratch array of the library (3.2 MiB), one copy
 * per GPU, held until the process ends.  Limits: at most 4096 scenes
 * (SPHE_ERANGE beyond); the context must already hold its device state (a
 * step or sphe_mass_all has run since the last edit; SPHE_ESTATE otherwise),
 * since building it allocates.  All calls on one GPU share the scratch:
 * calls that overlap on one GPU (several contexts on different threads or
 * streams) race and must be serialised by the caller. */
int sphe_mass_all_noalloc(sphe_ctx* ctx, double* terrain, double* sediment, uint64_t* units, int count);

/* Status word of the last sphe_step call, 0 when all went well:
 *   bit 1 (value 1): a non-finite position or velocity.  The reference
 *     produces NaN velocity for a particle at rest outside the box
 *     (fluid_system.h:345); this reports it.
 *   bit 2 (value 2): a particle the sort could not place (inconsistent
 *     counts); it was dropped, not written out of range.
 *   bit 3 (value 4): a particle's neighbours fitted neither its list nor a
 *     pool chunk, so the force pass walked its stencil again: the same
 *     result, slower (the lists grow for the next steps). */
int sphe_status(sphe_ctx* ctx, uint32_t* nonfinite);

/* ---- measurement ------------------------------------------------------- */

/* Per-kernel HIP event timing of sphe_step (graph capture is bypassed while
 * profiling is on). */
int sphe_profile(sphe_ctx* ctx, int enable);
/* Average ms and launch count of one kernel ("density", "force", ...) since
 * profiling was enabled; names lists them comma-separated. */
int sphe_profile_read(sphe_ctx* ctx, const char* kernel, double* avg_ms, int* launches);
int sphe_profile_names(sphe_ctx* ctx, char* names, size_t cap);
/* Algorithmic HBM bytes of one launch of a kernel for the current scenes. */
int sphe_kernel_bytes(sphe_ctx* ctx, const char* kernel, double* bytes);

/* ---- tuning ------------------------------------------------------------ */

/* Launch parameters: block size of the neighbour kernels (default 128) and
 * the density pass's XCD mapping: blocks go to the 8 XCDs in runs of
 * xcd_remap consecutive blocks (default 128: neighbouring cells share an
 * XCD's L2; 0 or 1: plain round-robin order).  The force pass always uses
 * the plain order, which keeps the XCDs balanced when neighbour counts vary
 * across the domain.  sphe_set_tuning sets the block size of both neighbour
 * passes; sphe_set_force_block then gives the force pass its own (0: the
 * density pass's), and sphe_get_force_block reads the one in effect.
 * sphe_autotune runs the step with block sizes {64, 128, 256, 512}, times
 * the density and force passes of each trial with the profiling events, and
 * keeps each pass's fastest, with the run as set.  No parameter here
 * changes results.  No counterpart in the reference (a serial CPU loop). */
int sphe_set_tuning(sphe_ctx* ctx, int block, int xcd_remap);
int sphe_get_tuning(const sphe_ctx* ctx, int* block, int* xcd_remap);
int sphe_set_force_block(sphe_ctx* ctx, int block);
int sphe_get_force_block(const sphe_ctx* ctx, int* block);
/* Lanes per particle in the density and force passes (1, 2, 4 or 8;
 * default 4).  Lane l takes positions l, l+G, ... of each stencil row and the
 * partial sums are merged in a fixed xor tree, so results are deterministic
 * for a given lane count (and differ between counts by summation order). */
int sphe_set_lanes(sphe_ctx* ctx, int lanes);
int sphe_get_lanes(const sphe_ctx* ctx, int* lanes);
/* x slabs per grid cell (1, 2 or 4; default 4).  A cell's particles are
 * ordered by (slab, id) and each stencil row is clipped to the slabs the
 * smoothing sphere reaches, so the neighbour passes test fewer candidates.
 * The slab count changes the canonical order, hence the summation order. */
int sphe_set_subcells(sphe_ctx* ctx, int slabs);
int sphe_get_subcells(const sphe_ctx* ctx, int* slabs);
/* Cell split along y and z (1, 2 or 3; default 2).  Cells stay h (1 + 2^-8)
 * along x and are 1/split of that along y and z, so a particle's stencil
 * has (2 split + 1)^2 rows, each clipped to the smoothing sphere: split 2
 * tests about 40% fewer candidates than split 1 for 2.7x the rows and 4x
 * the cells, split 3 about 16% fewer than split 2 for 1.9x the rows.  The
 * split changes the canonical order, hence the summation order; setting it
 * rebuilds the context's layout (the state is kept).  No counterpart in the
 * reference (which has no grid). */
int sphe_set_yz_split(sphe_ctx* ctx, int split);
int sphe_get_yz_split(const sphe_ctx* ctx, int* split);
/* Neighbour-list entries per particle (default 256; a multiple of 16).  A
 * particle with more neighbours continues its list in one chunk of an
 * overflow pool or, without room there, walks its stencil in the force pass
 * (status bit 4).  After a step sequence that ran out of room the lists grow
 * to 1.5x the largest count, as far as all lists together fit in a quarter
 * of the device's memory; longer neighbourhoods get pool chunks of 1.5x
 * their excess over the list, and the pool 1.5x the chunks the last step
 * claimed, as far as it fits in another quarter.  Setting the capacity sets
 * the chunk size too.  Neither ever changes results. */
int sphe_set_list_capacity(sphe_ctx* ctx, int entries);
int sphe_get_list_capacity(const sphe_ctx* ctx, int* entries);
/* Entries per overflow chunk and the pool's chunk count (0 before the
 * context has a device layout). */
int sphe_get_overflow_capacity(const sphe_ctx* ctx, int* chunk_entries, int64_t* chunks);
int sphe_autotune(sphe_ctx* ctx, int steps_per_trial);
/* Persistent density and force passes (default off): a grid of resident
 * blocks takes the passes' blocks of every scene from eight work queues,
 * one per XCD, each keeping the ordinary launch's XCD runs of cells, with
 * blocks of an empty queue taking work from the others.  Results are the
 * same bits either way; on MI355X the ordinary launches are faster (the
 * hardware refills a CU with no atomic per work item).  No counterpart in
 * the reference. */
int sphe_set_persistent(sphe_ctx* ctx, int on);
int sphe_get_persistent(const sphe_ctx* ctx, int* on);
/* Steps per captured graph launch of a long sphe_step call (1..4096,
 * default 32); the remaining steps of a call run as one-step graphs.  The
 * host reads the list-overflow words between launches, so longer sequences
 * grow the lists later.  Results do not depend on it.  No counterpart in
 * the reference (one Run per frame, fluid_system.h:104-183). */
int sphe_set_sequence_steps(sphe_ctx* ctx, int steps);
int sphe_get_sequence_steps(const sphe_ctx* ctx, int* steps);
/* Cells per scan tile of the sort's prefix sum: 2^log2_cells, 11..13
 * (default 12).  Results do not depend on it.  No counterpart in the
 * reference (which has no grid). */
int sphe_set_scan_tile(sphe_ctx* ctx, int log2_cells);
int sphe_get_scan_tile(const sphe_ctx* ctx, int* log2_cells);
/* Tunes the cell grid on the current state: every y/z split (1, 2) and
 * x-slab count (1, 2, 4), then every scan tile on the best of them, each
 * timed over steps_per_trial steps (all kernels, after two untimed steps)
 * from the same state; keeps the fastest and leaves the state as it was.
 * The split and the slab count set the canonical summation order, so the
 * tuned grid changes the last bits of later steps (every combination is a
 * tested configuration).  The choice is cached per (device, precision,
 * lanes, block sizes, scene sizes and smoothing lengths).  Run before
 * sphe_autotune, whose block sizes depend on the grid.  No counterpart in
 * the reference. */
int sphe_autotune_grid(sphe_ctx* ctx, int steps_per_trial);
/* Saves the particle and terrain state of every scene (a host copy) and
 * restores it later; stepping a restored state repeats the same steps bit
 * for bit.  No counterpart in the reference (whose state is the particle
 * vector itself); used to time and profile the same steps twice. */
int sphe_save_state(sphe_ctx* ctx);
int sphe_restore_state(sphe_ctx* ctx);

/* ---- verification hooks (used by the parity tests) ---------------------- */

/* Cell key of every particle (x fastest within the scene grid), id order. */
int sphe_debug_cell_keys(sphe_ctx* ctx, int scene, int32_t* keys);
/* Grid dimensions and derived float constants of a scene:
 * out = {h, h2, thr, t3h2, mc6, pf_scale, vf_scale, ng, cell, inv_cell,
 *        org x, org y, org z, dims x, dims y, dims z, hq, kcv, up, dn}. */
int sphe_debug_consts(const sphe_ctx* ctx, int scene, double* out, int cap);
/* Neighbour sets of the grid search (self included) from the current state:
 * counts[id order] and, if idx != NULL, the neighbour ids of each particle in
 * ascending id order at offsets[id order].  Pass idx == NULL to get counts. */
int sphe_debug_neighbours(sphe_ctx* ctx, int scene, int32_t* counts, const int64_t* offsets,
                          int32_t* idx);
/* Density of the current state (one density pass, no step), id order. */
int sphe_debug_density(sphe_ctx* ctx, int scene, double* rho);
/* Acceleration (F / rho, fluid_system.h:318-321) of every particle in the
 * current state, as the next step computes it (the same sums, bit for bit),
 * without stepping: acc[3 i + a] in id order.  The reference exposes it per
 * particle only, as FluidParticle::Acceleration through GetParticle
 * (fluid_system.h:286-289). */
int sphe_debug_accel(sphe_ctx* ctx, int scene, double* acc);
/* Plain device-to-device streaming copy of `bytes` (a multiple of 16), run
 * `iters` times on the context's stream: HBM bandwidth reference and the
 * known-size kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters.
 * Returns the average ms per copy. */
int sphe_debug_copy(sphe_ctx* ctx, size_t bytes, int iters, double* avg_ms);
/* Starts the overflow pool's chunk claims at entry `entries` (a multiple of
 * 16, at most half the pool's limit; 0 is the default) and allocates the
 * pool that much larger: a test can place chunks past 2^32 entries and check
 * that results do not change.  Rebuilds the device layout.  No counterpart in
 * the reference. */
int sphe_debug_pool_origin(sphe_ctx* ctx, int64_t entries);

#ifdef __cplusplus
}
#endif
#endif /* SPHE_H */
