// sphe_device.h -- scene descriptors and device helpers shared by the HIP
// kernels (sphe_kernels.hip) and the host runtime (sphe_runtime.cpp).
//
// Layout of one context in HBM (all scenes of a batch concatenated):
//   particles  slot ranges [p_begin, p_begin + p_cap) per scene, p_cap a
//              multiple of 64 (one wavefront); the first n slots of a scene
//              hold particles, the rest is padding that sorts into the
//              scene's pad cell and is skipped by every physics kernel.
//   cells      dense uniform grid per scene, x fastest, global keys
//              [cell_off, cell_off + ncell) plus one pad cell at
//              cell_off + ncell; start[] is the exclusive prefix of counts.
//   terrain    cells [t_off, t_off + tcap) per scene, row = z.
#pragma once

#include <stdint.h>

#define SPHE_WAVE 64
// cells per scan tile: 2^scan_log2 (256 threads x 8, 16 or 32 counts)
// Overflow-pool chunk offsets are stored in units of this many entries, so
// 32-bit offsets reach 2^36 entries.
#define SPHE_POOL_UNIT 16

#define SPHE_SCAN_LOG2_MIN 11
#define SPHE_SCAN_LOG2_DEFAULT 12
#define SPHE_SCAN_LOG2_MAX 13

template <typename T>
struct SceneConsts {
    T h, h2, thr, t3h2, eps_r, inv_sqrt3;
    T mc6;       // mass * poly6 constant 315 / (64 pi h^9)
    T pf_scale;  // mass * spiky-gradient constant 45 / (pi h^6)
    T vf_scale;  // visc * viscosity-laplacian constant 45 / (pi h^6)
    T ng;        // -(poly6 gradient constant 945 / (32 pi h^9))
    T rho0, k, mass, surf_tens, g[3], dt;
    T box_c[3], box_lo[3], box_hi[3];
    T inv_cell, org[3];
    T inv_yz;    // inv_cell * yz: y and z cell coordinates
    T yz_scale;  // 1 / yz: y/z cell units to x-cell units
    T t_org[3], t_inv, kcv, ke, kd, band, up, dn, dmax;
};

struct SceneDev {
    int32_t p_begin, p_cap, n, tcap;
    int32_t cell_off, ncell, dims[3];
    int32_t t_off, tnx, tnz, erosion, fx_exp;
    float hq;    // height = t_org[1] + (float)vol * hq
    float t_y0;
    // id -> sorted slot map of the last step (sphe_get_particle): entries
    // [idmap_off, idmap_off + idmap_len) for ids 0 .. idmap_len - 1; -1: the
    // scene's ids are too sparse for a map and are searched instead
    int32_t idmap_off, idmap_len;
    SceneConsts<float> f;
    SceneConsts<double> d;
};

template <typename T>
struct SceneSel;
template <>
struct SceneSel<float> {
    static __host__ __device__ inline const SceneConsts<float>& get(const SceneDev& s) { return s.f; }
};
template <>
struct SceneSel<double> {
    static __host__ __device__ inline const SceneConsts<double>& get(const SceneDev& s) { return s.d; }
};

// Device pointers and sizes of one context, passed by value to launchers.
struct DevState {
    int64_t P;          // total particle slots (multiple of 64)
    int64_t C;          // total cells including one pad cell per scene
    int64_t TC;         // total terrain cells (multiple of 64 per scene)
    int nscenes;
    int precision;      // SPHE_FP32 / SPHE_FP32_NAIVE / SPHE_FP64
    const SceneDev* scenes;
    const int32_t* chunk_scene;   // scene of each 64-slot particle chunk
    const int32_t* tchunk_scene;  // scene of each 64-cell terrain chunk
    // state in unsorted slot order (between steps) and sorted order (in a step)
    void* u_pos; void* u_vel; int64_t* u_sed; int32_t* u_id;
    void* s_pos; void* s_vel; int64_t* s_sed; int32_t* s_id;
    void* rho; void* posa;              // density; (x, y, z, p / rho^2) per sorted slot
    int32_t* key; uint32_t* arank;      // cell key << sub_log2 | x slab, arrival rank in the cell
    uint64_t* tmp_kid; int32_t* tmp_src;  // scatter output: (key, id) packed, previous slot
    uint32_t* cnt; uint32_t* start; uint32_t* partial; int64_t npartial;
    // one word per scan tile of cnt (2^scan_log2 cells): set when a count
    // in the tile is taken, cleared by the scan that consumes the tile; the
    // scan reads the counts of flagged tiles only (the rest are zero)
    uint32_t* tocc;
    // one word per scan tile of start: the value every start of the tile
    // holds when the tile was last written empty, ~0 otherwise; the scan
    // leaves an empty tile unwritten when that value is still right
    uint32_t* tfill;
    int32_t scan_log2;  // cells per scan tile: 2^scan_log2
    // cells are split into 2^sub_log2 slabs along x (sub_log2 <= 2); subw
    // holds each cell's slab offsets (10 bits per inner boundary; bit 31:
    // more than 1023 particles, no offsets), rewritten every step for the
    // occupied cells (an empty cell's word may be stale: readers clamp the
    // offsets to the cell)
    uint32_t* subw; int32_t sub_log2;
    // cells are 1/yz of the x edge along y and z (yz = 1, 2 or 3): the stencil
    // has (2 yz + 1)^2 rows
    int32_t yz;
    int64_t* vol; float* height;
    int64_t* dq; int32_t* tcell;        // two-phase erosion records
    uint32_t* tcnt; uint32_t* tstart; uint32_t* tpartial; int64_t ntpartial;
    int64_t* tdq;                       // two-phase: deltas grouped by terrain cell
    uint32_t* status;
    double* red;                        // reduction scratch
    int64_t nred;
    // neighbour lists written by the density pass for the force pass: the
    // neighbours of sorted slot i (self excluded) in canonical order at
    // nbl[i * maxn + m], m < min(nbc[i], maxn); a particle whose count
    // exceeds maxn walks its stencil again in the force pass
    uint32_t* nbl;
    uint32_t* nbc;
    int32_t maxn;
    // overflow pool: a particle whose count passes maxn claims one chunk of
    // ocap more entries (pool[16 pbase[i] .. + ocap), neighbours maxn ..
    // maxn + ocap - 1) with an atomic on *ptop, which the scan of every step
    // resets to pool_origin; pbase[i] is the chunk's first entry in units of
    // SPHE_POOL_UNIT entries (ocap and pool_origin are multiples of it), or
    // ~0u when it has none (the force pass then walks its stencil beyond
    // maxn and sets status bit 4).  pool_cap counts entries from pool[0].
    uint32_t* pool;
    unsigned long long* ptop;
    uint32_t* pbase;
    int64_t pool_cap;
    int64_t pool_origin;
    int32_t ocap;
    // ovf[0]: largest neighbour count that overflowed a list; ovf[1]:
    // particles that found the pool full (both since the host last grew them)
    uint32_t* ovf;
    // persistent density and force passes (when set): a resident grid takes
    // the passes' blocks from eight queues, work[0..7]
    int32_t persistent;
    uint32_t* work;
};
