// sphe_runtime.cpp -- host runtime and C ABI (include/sphe.h) of the MI355X
// SPH-Erosion step.
//
// A context owns one HIP stream and the device buffers of all its scenes
// (layout in sphe_device.h).  Host staging arrays keep each scene's particles
// in ascending id order, as the reference's std::vector<FluidParticle> keeps
// them in insertion order (Erosion/fluid_system.h:482); they are the source
// of truth after a host-side edit (Initialize, AddParticles, Reset,
// set_particles, terrain load) and are refreshed from the device on demand
// after steps.  The device state is rebuilt from them whenever the particle
// counts or the grid change.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/sphe.h"
#include "sphe_device.h"
#include "sphe_launch.h"

namespace {

const float REF_PI = 3.141592f;  // Erosion/sphere.h:8

struct Scene {
    sphe_params prm;
    float origin[3] = {0, 0, 0};  // m_Origin (fluid_system.h:484)
    int init_num = 0;             // fluid_system.h:474
    int next_id = 0;              // fluid_system.h:479
    // terrain
    int tnx = 0, tnz = 0;
    float t_org[3] = {0, 0, 0};
    float t_cell = 1.0f;
    int fx_exp = 60;
    std::vector<int64_t> vol;  // host copy of the fixed-point volumes
    // particles in ascending id order
    std::vector<double> pos, vel;
    std::vector<int32_t> ids;
    std::vector<int64_t> sed;
    // layout (valid while the context layout is)
    int64_t p_begin = 0, p_cap = 0, cell_off = 0, ncell = 0, t_off = 0, tcap = 0;
    int dims[3] = {0, 0, 0};
    int64_t idmap_off = -1, idmap_len = 0;  // the scene's part of the id -> slot map (-1: none)
    bool ids_dense = false;                 // the ids are 0 .. n-1 (the device gathers in id order)
};

struct ProfRec {
    std::string name;
    hipEvent_t a, b;
};

}  // namespace

struct sphe_ctx {
    int precision = SPHE_FP32;
    int device = 0;
    std::string err;
    hipStream_t stream = nullptr;
    // a deque: adding a scene never moves the others, so the parameter
    // storage sphe_param_ptr hands out stays where it is
    std::deque<Scene> scenes;

    bool layout_valid = false;
    bool device_valid = false;  // device holds the latest state
    bool host_valid = true;     // host staging holds the latest state
    bool counted = false;       // key / arank / cnt prepared for the next scan
    bool heights_stale = false; // the last step eroded; heights follow in the next step's scan
    bool stepped = false;       // sorted buffers hold the last step's pre-step state
    // state epoch: bumped by every call that changes the device state or how
    // it is read; the verification hooks' sort of the current state is kept
    // while its epoch is current
    uint64_t epoch = 1;
    uint64_t sorted_epoch = 0;

    DevState D{};
    std::vector<void*> allocs;
    std::vector<SceneDev> table;  // host copy of the uploaded scene table
    bool table_uploaded = false;

    bool use_graph = true;
    // captured step sequences by length (steps per graph launch)
    std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
    std::string graph_sig;
    // the list-overflow word as the last step sequence published it: host
    // memory mapped for the device (h_ovf_dev), written by a kernel at the
    // end of every sequence and read by the host without a stream wait
    uint32_t* h_ovf = nullptr;
    uint32_t* h_ovf_dev = nullptr;
    // completion of the last launches (one event per launch slot): before a
    // launch the host waits for the one kInFlight launches back, so the
    // overflow word it reads is at most that many launches old
    hipEvent_t seq_ev[2] = {nullptr, nullptr};
    uint64_t seq_count = 0;

    bool profiling = false;
    std::vector<ProfRec> pending;
    std::vector<hipEvent_t> event_pool;
    std::map<std::string, std::pair<double, int>> prof;
    std::map<std::pair<int, int>, std::vector<float>> views;  // sphe_buffer snapshots by (scene, buffer)

    int block = 128;
    int fblock = 0;  // block size of the force pass (0: the density pass's)
    int remap = 128;  // density: runs of 128 consecutive blocks per XCD (L2 locality); force: plain order
    int maxn = 256;  // neighbour-list capacity per particle (grown when a step overflows it)
    int ocap = 256;  // entries per overflow-pool chunk
    size_t mem_total = 0;  // the device's memory (bounds the neighbour lists)
    int64_t pool_chunks = 0;  // overflow chunks the last growth asked for (0: none yet)
    int64_t pool_origin = 0;  // first pool entry the chunk claims start from (a test hook; 0)
    int seq_steps = 32;       // steps per captured graph launch of a long sphe_step call
    int lanes = 4;   // lanes per particle in the density and force passes
    int scan_log2 = SPHE_SCAN_LOG2_DEFAULT;  // cells per scan tile: 2^scan_log2
    bool persistent = false;  // density and force as persistent launches over work queues
    int sub_log2 = 2;  // x slabs per cell: 2^sub_log2
    int yz = 2;        // cells are 1/yz of the x edge along y and z

    double* d_mass_part = nullptr;
    double* d_mass_out = nullptr;
    // sphe_mass_all: per-scene block partials, then the per-scene results
    double* d_mass_all = nullptr;
    size_t mass_all_cap = 0;  // doubles
    // the library's static scratch on this context's device (the device
    // address of its copy), for sphe_mass_all_noalloc; null if the lookup failed
    double* noalloc_scratch = nullptr;
    size_t noalloc_doubles = 0;
    double* d_inspect = nullptr;
    // id -> sorted slot map of the last step (sphe_get_particle), rebuilt on
    // the first inspection after a step; d_slot: the slot of a searched id
    int32_t* d_idmap = nullptr;
    int64_t idmap_len = 0;
    bool idmap_valid = false;
    int32_t* d_slot = nullptr;
    // lane count and slab count the last step ran with: its inspection
    // repeats its sums in the same order whatever was set since
    int step_lanes = 4;
    int step_sub_log2 = 2;
    float* d_mesh = nullptr;  // terrain vertex stream (sphe_terrain_mesh*)
    size_t mesh_cap = 0;      // floats
    float* d_feed = nullptr;  // one scene's positions or velocities in id order
    size_t feed_cap = 0;      // floats

    // sphe_save_state / sphe_restore_state: particle and terrain state of
    // every scene, in the host staging layout
    struct Saved {
        std::vector<double> pos, vel;
        std::vector<int32_t> ids;
        std::vector<int64_t> sed, vol;
        int next_id;
    };
    std::vector<Saved> saved;
    bool has_saved = false;
};

namespace {

int fail(sphe_ctx* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK(ctx, expr)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail((ctx), SPHE_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// hipMalloc that leaves no error behind when it fails: the runtime keeps a
// failed call's error as its last error, and the launch checks that follow
// (hipGetLastError) would report it as a failure of their own.
template <typename T>
hipError_t try_malloc(T** p, size_t bytes)
{
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        q = nullptr;
    }
    *p = (T*)q;
    return e;
}

size_t real_size(int precision) { return precision == SPHE_FP64 ? sizeof(double) : sizeof(float); }

// Largest overflow pool in entries, from pool[0]: a quarter of the device's
// memory, and chunk offsets that fit 32 bits in units of SPHE_POOL_UNIT
// entries (below the ~0u that means no chunk).
int64_t pool_max(const sphe_ctx* c)
{
    const int64_t by_offset = ((int64_t)UINT32_MAX - 1) * SPHE_POOL_UNIT;
    return std::min<int64_t>(by_offset, (int64_t)(c->mem_total / 4) / 4);
}

// Overflow pool entries for `chunks` chunks of ocap entries after the
// context's pool origin.
int64_t pool_entries(const sphe_ctx* c, int64_t chunks, int ocap)
{
    return std::min<int64_t>(c->pool_origin + chunks * ocap, pool_max(c));
}

// Overflow chunks of a new layout before any step has claimed one: one for
// an eighth of the particle slots, but at most 16 entries per slot in
// chunks of ocap entries (at least 64 chunks).
int64_t pool_chunks_initial(int64_t P, int ocap)
{
    return std::max<int64_t>(std::min<int64_t>(P / 8, P * 16 / ocap), 64);
}

// ---------------------------------------------------------------- derive

float pred_threshold(float h)
{
    // glm::length(r) <= h (fluid_system.h:118-119) <=> d2 <= T(h), the
    // largest float whose correctly rounded square root is <= h.
    float t = h * h;
    while (std::sqrt(std::nextafter(t, INFINITY)) <= h) t = std::nextafter(t, INFINITY);
    while (t > 0.0f && std::sqrt(t) > h) t = std::nextafter(t, 0.0f);
    return t;
}

// Cells are cell = h (1 + 2^-8) along x (so 3 cells always cover the
// radius under FP32 rounding) and cell / yz along y and z (2 yz + 1 cells).
void grid_dims(const sphe_params& p, int dims[3], int yz)
{
    const float h = p.h;
    const float cell = h * 1.00390625f;
    const float inv = 1.0f / cell;
    for (int a = 0; a < 3; ++a) {
        const float lo = p.box_center[a] - p.box_half;
        const float hi = p.box_center[a] + p.box_half;
        const float ia = a == 0 ? inv : inv * (float)yz;  // exact for yz 1 and 2
        dims[a] = (int)std::floor((hi - lo) * ia) + 1;
    }
}

void derive(const Scene& s, SceneDev& d, int yz)
{
    const sphe_params& p = s.prm;
    SceneConsts<float>& f = d.f;
    const float h = p.h;
    f.h = h;
    f.h2 = h * h;
    f.thr = pred_threshold(h);
    f.t3h2 = 3.0f * f.h2;
    // dist < 10e-5 (fluid_system.h:438): float dist against the double 1e-4,
    // i.e. dist <= fl32(1e-4)
    f.eps_r = std::nextafter(1e-4f, INFINITY);
    f.inv_sqrt3 = 1.0f / std::sqrt(3.0f);
    const double PI_D = (double)REF_PI;
    const double hd = (double)h;
    const double h6 = hd * hd * hd * hd * hd * hd, h9 = h6 * hd * hd * hd;
    const double mass = (double)p.mass, visc = (double)p.visc;
    f.mc6 = (float)(mass * (315.0 / (64.0 * PI_D * h9)));
    f.pf_scale = (float)(mass * (45.0 / (PI_D * h6)));
    f.vf_scale = (float)(visc * (45.0 / (PI_D * h6)));
    f.ng = (float)(-(945.0 / (32.0 * PI_D * h9)));
    f.rho0 = p.rest_density;
    f.k = p.stiffness;
    f.mass = p.mass;
    f.surf_tens = p.surf_tens;
    f.dt = p.dt;
    const float cell = h * 1.00390625f;
    f.inv_cell = 1.0f / cell;
    f.inv_yz = f.inv_cell * (float)yz;
    f.yz_scale = 1.0f / (float)yz;
    for (int a = 0; a < 3; ++a) {
        f.g[a] = p.gravity[a];
        f.box_c[a] = p.box_center[a];
        f.box_lo[a] = p.box_center[a] - p.box_half;
        f.box_hi[a] = p.box_center[a] + p.box_half;
        f.org[a] = f.box_lo[a];
        f.t_org[a] = s.t_org[a];
    }
    grid_dims(p, d.dims, yz);
    f.t_inv = 1.0f / s.t_cell;
    const double area = (double)s.t_cell * (double)s.t_cell;
    const double kc = (double)p.kc, rho0 = (double)p.rest_density;
    f.kcv = (float)(kc * mass / rho0);
    f.ke = p.ke;
    f.kd = p.kd;
    f.band = p.band;
    f.up = (float)std::ldexp(1.0, s.fx_exp);
    f.dn = (float)std::ldexp(1.0, -s.fx_exp);
    f.dmax = (float)std::ldexp(1.0, 52 - s.fx_exp);

    // FP64 comparison path: the same formulas in double from the FP32 values
    SceneConsts<double>& g = d.d;
    g.h = hd;
    g.h2 = hd * hd;
    g.thr = hd * hd;
    g.t3h2 = 3.0 * hd * hd;
    g.eps_r = 1e-4;
    g.inv_sqrt3 = 1.0 / std::sqrt(3.0);
    g.mc6 = mass * (315.0 / (64.0 * PI_D * h9));
    g.pf_scale = mass * (45.0 / (PI_D * h6));
    g.vf_scale = visc * (45.0 / (PI_D * h6));
    g.ng = -(945.0 / (32.0 * PI_D * h9));
    g.rho0 = rho0;
    g.k = (double)p.stiffness;
    g.mass = mass;
    g.surf_tens = (double)p.surf_tens;
    g.dt = (double)p.dt;
    g.inv_cell = 1.0 / (hd * (1.0 + std::ldexp(1.0, -8)));
    g.inv_yz = g.inv_cell * (double)yz;
    g.yz_scale = 1.0 / (double)yz;
    for (int a = 0; a < 3; ++a) {
        g.g[a] = (double)p.gravity[a];
        g.box_c[a] = (double)f.box_c[a];
        g.box_lo[a] = (double)f.box_lo[a];
        g.box_hi[a] = (double)f.box_hi[a];
        g.org[a] = (double)p.box_center[a] - (double)p.box_half;
        g.t_org[a] = (double)s.t_org[a];
    }
    g.t_inv = 1.0 / (double)s.t_cell;
    g.kcv = kc * mass / rho0;
    g.ke = (double)p.ke;
    g.kd = (double)p.kd;
    g.band = (double)p.band;
    g.up = std::ldexp(1.0, s.fx_exp);
    g.dn = std::ldexp(1.0, -s.fx_exp);
    g.dmax = std::ldexp(1.0, 52 - s.fx_exp);

    d.hq = (float)(std::ldexp(1.0, -s.fx_exp) / area);
    d.t_y0 = s.t_org[1];
    d.tnx = s.tnx;
    d.tnz = s.tnz;
    d.fx_exp = s.fx_exp;
    d.erosion = (s.tnx > 0 && s.tnz > 0) ? p.erosion : 0;
    d.p_begin = (int32_t)s.p_begin;
    d.p_cap = (int32_t)s.p_cap;
    d.n = (int32_t)s.ids.size();
    d.tcap = (int32_t)s.tcap;
    d.cell_off = (int32_t)s.cell_off;
    d.ncell = (int32_t)s.ncell;
    d.t_off = (int32_t)s.t_off;
    d.idmap_off = (int32_t)s.idmap_off;
    d.idmap_len = (int32_t)s.idmap_len;
}

// ---------------------------------------------------------------- lattice

// FluidSystemSPH::Initialize / AddParticles lattice (fluid_system.h:80-95,
// 234-249): i, j, k < cbrt(n); coordinates computed in double, narrowed to
// float, origin added in float.
void append_lattice(Scene& s, int n)
{
    const double cb = std::cbrt((double)n);
    int side = 0;
    while ((double)side < cb) ++side;
    for (int i = 0; i < side; ++i)
        for (int j = 0; j < side; ++j)
            for (int k = 0; k < side; ++k) {
                const float x = (float)(-0.2 + i * 0.025);
                const float y = (float)(-0.05 + j * 0.025);
                const float z = (float)(-0.15 + k * 0.025);
                s.pos.push_back((double)(x + s.origin[0]));
                s.pos.push_back((double)(y + s.origin[1]));
                s.pos.push_back((double)(z + s.origin[2]));
                for (int a = 0; a < 3; ++a) s.vel.push_back(0.0);
                s.ids.push_back(s.next_id++);
                s.sed.push_back(0);
            }
}

// keeps each scene's staging arrays in ascending id order (stable)
void sort_by_id(Scene& s)
{
    const size_t n = s.ids.size();
    bool sorted = true;
    for (size_t i = 1; i < n && sorted; ++i) sorted = s.ids[i - 1] <= s.ids[i];
    if (sorted) return;
    std::vector<size_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return s.ids[a] < s.ids[b]; });
    std::vector<double> p(3 * n), v(3 * n);
    std::vector<int32_t> id(n);
    std::vector<int64_t> sd(n);
    for (size_t t = 0; t < n; ++t) {
        const size_t i = perm[t];
        for (int a = 0; a < 3; ++a) {
            p[3 * t + a] = s.pos[3 * i + a];
            v[3 * t + a] = s.vel[3 * i + a];
        }
        id[t] = s.ids[i];
        sd[t] = s.sed[i];
    }
    s.pos.swap(p);
    s.vel.swap(v);
    s.ids.swap(id);
    s.sed.swap(sd);
}

// ------------------------------------------------------------- buffers

void drop_graphs(sphe_ctx* c)
{
    for (auto& kv : c->graphs) {
        if (kv.second.second) (void)hipGraphExecDestroy(kv.second.second);
        if (kv.second.first) (void)hipGraphDestroy(kv.second.first);
    }
    c->graphs.clear();
    c->graph_sig.clear();
}

void free_device(sphe_ctx* c)
{
    drop_graphs(c);
    for (void* p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    c->D = DevState{};
    c->layout_valid = false;
    c->device_valid = false;
    c->counted = false;
    c->stepped = false;
    c->table_uploaded = false;
    c->heights_stale = false;
    ++c->epoch;
    c->d_idmap = nullptr;
    c->idmap_len = 0;
    c->idmap_valid = false;
}

int dalloc(sphe_ctx* c, void** p, size_t bytes)
{
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = try_malloc(p, bytes);
    if (e != hipSuccess) return fail(c, SPHE_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    c->allocs.push_back(*p);
    return SPHE_OK;
}

#define DALLOC(ptr, bytes)                                        \
    do {                                                          \
        void* p_;                                                 \
        int r_ = dalloc(c, &p_, (bytes));                         \
        if (r_) return r_;                                        \
        ptr = (decltype(ptr))p_;                                  \
    } while (0)

// Zeroes the four published words (largest overflowing count, failed pool
// claims, pool entries claimed) in the mapped host block.
void clear_published(sphe_ctx* c)
{
    volatile uint32_t* h = c->h_ovf;
    for (int k = 0; k < 4; ++k) h[k] = 0u;
}

int build_layout(sphe_ctx* c)
{
    free_device(c);
    int64_t P = 0, C = 0, TC = 0;
    for (Scene& s : c->scenes) {
        s.p_begin = P;
        s.p_cap = round_up((int64_t)s.ids.size(), SPHE_WAVE);
        P += s.p_cap;
        grid_dims(s.prm, s.dims, c->yz);
        // the neighbour passes clamp cell and x sub-cell indices (4 per cell
        // at most) in FP32, exact below 2^24
        if ((int64_t)s.dims[0] * 4 > (int64_t)1 << 24 || s.dims[1] > 1 << 24 || s.dims[2] > 1 << 24)
            return fail(c, SPHE_ERANGE, "grid too long (%d x %d x %d cells)", s.dims[0], s.dims[1], s.dims[2]);
        s.ncell = (int64_t)s.dims[0] * s.dims[1] * s.dims[2];
        s.cell_off = C;
        C += s.ncell + 1;
        s.t_off = TC;
        s.tcap = round_up((int64_t)s.tnx * s.tnz, SPHE_WAVE);
        TC += s.tcap;
    }
    // id -> slot map: one entry per id 0 .. max id of each scene whose ids
    // are dense enough (max id below 2 n + 4096; ids are sorted ascending in
    // the staging arrays); the others are searched on inspection
    int64_t M = 0;
    for (Scene& s : c->scenes) {
        s.idmap_off = -1;
        s.idmap_len = 0;
        s.ids_dense = true;
        for (size_t i = 0; i < s.ids.size() && s.ids_dense; ++i) s.ids_dense = s.ids[i] == (int32_t)i;
        if (s.ids.empty()) continue;
        const int64_t len = (int64_t)s.ids.back() + 1;
        if (len > 2 * (int64_t)s.ids.size() + 4096 || M + len > INT32_MAX / 2) continue;
        s.idmap_off = M;
        s.idmap_len = len;
        M += len;
    }
    // per-slot vectors are addressed with 32-bit byte offsets: each stays under
    // 4 GiB, with room for the density pass's positions a block past the end
    const int64_t vec_bytes = 4 * (int64_t)real_size(c->precision);
    if (P > INT32_MAX / 2 || C > INT32_MAX / 2 || TC > INT32_MAX / 2 ||
        (P + 64) * vec_bytes > (int64_t)UINT32_MAX)
        return fail(c, SPHE_ERANGE, "context too large (slots %lld, cells %lld)", (long long)P, (long long)C);
    DevState& D = c->D;
    D.P = P;
    D.C = C;
    D.TC = TC;
    D.nscenes = (int)c->scenes.size();
    D.precision = c->precision;
    const size_t E = real_size(c->precision);
    const size_t Pn = (size_t)P;
    DALLOC(D.u_pos, Pn * 4 * E);
    DALLOC(D.u_vel, Pn * 4 * E);
    DALLOC(D.s_pos, Pn * 4 * E);
    DALLOC(D.s_vel, Pn * 4 * E);
    DALLOC(D.u_sed, Pn * 8);
    DALLOC(D.s_sed, Pn * 8);
    DALLOC(D.u_id, Pn * 4);
    DALLOC(D.s_id, Pn * 4);
    DALLOC(D.rho, Pn * E);
    DALLOC(D.posa, Pn * 4 * E);
    DALLOC(D.key, Pn * 4);
    DALLOC(D.arank, Pn * 4);
    DALLOC(D.tmp_kid, Pn * 8);
    DALLOC(D.tmp_src, Pn * 4);
    DALLOC(D.cnt, (size_t)(C + 1) * 4);
    DALLOC(D.start, (size_t)(C + 1) * 4);
    DALLOC(D.subw, (size_t)(C + 1) * 4);
    D.yz = c->yz;
    D.sub_log2 = c->sub_log2;
    D.scan_log2 = c->scan_log2;
    while (D.sub_log2 > 0 && (C << D.sub_log2) > INT32_MAX) --D.sub_log2;  // keys are int32
    D.npartial = sphe_scan_partials(C + 1);
    DALLOC(D.partial, (size_t)D.npartial * 4);
    DALLOC(D.tocc, (size_t)D.npartial * 4);
    DALLOC(D.tfill, (size_t)D.npartial * 4);
    DALLOC(D.vol, (size_t)TC * 8);
    DALLOC(D.height, (size_t)TC * 4);
    DALLOC(D.dq, Pn * 8);
    DALLOC(D.tcell, Pn * 4);
    DALLOC(D.tcnt, (size_t)(TC + 1) * 4);
    DALLOC(D.tstart, (size_t)(TC + 1) * 4);
    D.ntpartial = sphe_scan_partials(TC + 1);
    DALLOC(D.tpartial, (size_t)D.ntpartial * 4);
    DALLOC(D.tdq, Pn * 8);
    D.maxn = c->maxn;
    DALLOC(D.nbl, Pn * (size_t)D.maxn * 4);
    DALLOC(D.nbc, Pn * 4);
    D.ocap = c->ocap;
    // a rebuilt layout keeps the pool the steps before it sized
    D.pool_origin = c->pool_origin;
    D.pool_cap =
        pool_entries(c, c->pool_chunks > 0 ? c->pool_chunks : pool_chunks_initial(P, D.ocap), D.ocap);
    DALLOC(D.pool, (size_t)D.pool_cap * 4);
    DALLOC(D.pbase, Pn * 4);
    DALLOC(D.ptop, 64);
    DALLOC(D.ovf, 64);
    DALLOC(D.status, 64);
    DALLOC(D.work, 64);
    D.persistent = c->persistent ? 1 : 0;
    DALLOC(c->d_mass_part, (size_t)sphe_mass_blocks() * 3 * 8);
    DALLOC(c->d_mass_out, 64);
    DALLOC(c->d_inspect, 64 * 8);
    DALLOC(c->d_slot, 64);
    if (M > 0) DALLOC(c->d_idmap, (size_t)M * 4);
    c->idmap_len = M;
    c->idmap_valid = false;
    SceneDev* sc;
    DALLOC(sc, sizeof(SceneDev) * std::max<size_t>(1, c->scenes.size()));
    D.scenes = sc;
    int32_t* cs;
    DALLOC(cs, (size_t)(P / SPHE_WAVE + 1) * 4);
    D.chunk_scene = cs;
    int32_t* tcs;
    DALLOC(tcs, (size_t)(TC / SPHE_WAVE + 1) * 4);
    D.tchunk_scene = tcs;

    std::vector<int32_t> hcs((size_t)(P / SPHE_WAVE + 1), 0), htcs((size_t)(TC / SPHE_WAVE + 1), 0);
    for (size_t si = 0; si < c->scenes.size(); ++si) {
        const Scene& s = c->scenes[si];
        for (int64_t k = s.p_begin / SPHE_WAVE; k < (s.p_begin + s.p_cap) / SPHE_WAVE; ++k) hcs[k] = (int32_t)si;
        for (int64_t k = s.t_off / SPHE_WAVE; k < (s.t_off + s.tcap) / SPHE_WAVE; ++k) htcs[k] = (int32_t)si;
    }
    HIPCHK(c, hipMemcpyAsync(cs, hcs.data(), hcs.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(tcs, htcs.data(), htcs.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemsetAsync(D.cnt, 0, (size_t)(C + 1) * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(D.tocc, 0, (size_t)D.npartial * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(D.tfill, 0xFF, (size_t)D.npartial * 4, c->stream));  // no tile known filled
    HIPCHK(c, hipMemsetAsync(D.subw, 0, (size_t)(C + 1) * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(D.tcnt, 0, (size_t)(TC + 1) * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(D.status, 0, 64, c->stream));
    HIPCHK(c, hipMemsetAsync(D.ovf, 0, 64, c->stream));
    HIPCHK(c, hipMemsetAsync(D.ptop, 0, 64, c->stream));
    if (TC > 0) HIPCHK(c, hipMemsetAsync(D.vol, 0, (size_t)TC * 8, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    clear_published(c);  // no overflow published for the new lists
    c->layout_valid = true;
    c->device_valid = false;
    return SPHE_OK;
}

template <typename T>
void pack_particles(const Scene& s, int64_t cap, std::vector<T>& pos4, std::vector<T>& vel4,
                    std::vector<int64_t>& sed, std::vector<int32_t>& ids)
{
    const int64_t n = (int64_t)s.ids.size();
    pos4.assign((size_t)cap * 4, T(0));
    vel4.assign((size_t)cap * 4, T(0));
    sed.assign((size_t)cap, 0);
    ids.assign((size_t)cap, 0);
    for (int64_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            pos4[4 * i + a] = (T)s.pos[3 * i + a];
            vel4[4 * i + a] = (T)s.vel[3 * i + a];
        }
        sed[i] = s.sed[i];
        ids[i] = s.ids[i];
    }
    // padding: unique ids above the scene's largest, parked at the box centre
    int32_t top = 0;
    for (int64_t i = 0; i < n; ++i) top = std::max(top, s.ids[i]);
    for (int64_t i = n; i < cap; ++i) {
        for (int a = 0; a < 3; ++a) pos4[4 * i + a] = (T)s.prm.box_center[a];
        ids[i] = top + 1 + (int32_t)(i - n);
    }
}

template <typename T>
int upload_t(sphe_ctx* c)
{
    DevState& D = c->D;
    for (const Scene& s : c->scenes) {
        if (s.p_cap == 0) continue;
        std::vector<T> p4, v4;
        std::vector<int64_t> sd;
        std::vector<int32_t> id;
        pack_particles<T>(s, s.p_cap, p4, v4, sd, id);
        const size_t off = (size_t)s.p_begin;
        HIPCHK(c, hipMemcpy((T*)D.u_pos + 4 * off, p4.data(), p4.size() * sizeof(T), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy((T*)D.u_vel + 4 * off, v4.data(), v4.size() * sizeof(T), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(D.u_sed + off, sd.data(), sd.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(D.u_id + off, id.data(), id.size() * 4, hipMemcpyHostToDevice));
    }
    return SPHE_OK;
}

int upload_table(sphe_ctx* c, bool force);

int upload(sphe_ctx* c)
{
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int r = c->precision == SPHE_FP64 ? upload_t<double>(c) : upload_t<float>(c);
    if (r) return r;
    DevState& D = c->D;
    for (const Scene& s : c->scenes) {
        const size_t nt = (size_t)s.tnx * s.tnz;
        if (nt == 0) continue;
        HIPCHK(c, hipMemcpy(D.vol + s.t_off, s.vol.data(), nt * 8, hipMemcpyHostToDevice));
    }
    r = upload_table(c, true);
    if (r) return r;
    sphe_launch_heights(D, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->heights_stale = false;
    c->device_valid = true;
    c->counted = false;
    c->stepped = false;
    ++c->epoch;
    return SPHE_OK;
}

template <typename T>
int download_t(sphe_ctx* c)
{
    DevState& D = c->D;
    for (Scene& s : c->scenes) {
        const size_t n = s.ids.size();
        if (n == 0) continue;
        std::vector<T> p4(4 * n), v4(4 * n);
        std::vector<int64_t> sd(n);
        std::vector<int32_t> id(n);
        const size_t off = (size_t)s.p_begin;
        HIPCHK(c, hipMemcpy(p4.data(), (T*)D.u_pos + 4 * off, p4.size() * sizeof(T), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(v4.data(), (T*)D.u_vel + 4 * off, v4.size() * sizeof(T), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(sd.data(), D.u_sed + off, n * 8, hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(id.data(), D.u_id + off, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            for (int a = 0; a < 3; ++a) {
                s.pos[3 * i + a] = (double)p4[4 * i + a];
                s.vel[3 * i + a] = (double)v4[4 * i + a];
            }
            s.ids[i] = id[i];
            s.sed[i] = sd[i];
        }
        sort_by_id(s);
    }
    return SPHE_OK;
}

int sync_to_host(sphe_ctx* c)
{
    if (c->host_valid) return SPHE_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int r = c->precision == SPHE_FP64 ? download_t<double>(c) : download_t<float>(c);
    if (r) return r;
    for (Scene& s : c->scenes) {
        const size_t nt = (size_t)s.tnx * s.tnz;
        if (nt == 0) continue;
        HIPCHK(c, hipMemcpy(s.vol.data(), c->D.vol + s.t_off, nt * 8, hipMemcpyDeviceToHost));
    }
    c->host_valid = true;
    return SPHE_OK;
}

// Called before a host-side edit of particles, terrain or the grid.
int begin_host_edit(sphe_ctx* c)
{
    int r = sync_to_host(c);
    if (r) return r;
    c->device_valid = false;
    c->stepped = false;
    c->layout_valid = false;
    ++c->epoch;
    return SPHE_OK;
}

// Heights of the volumes the last step eroded (the next step's scan would
// compute them), for readers between steps.
int ensure_heights(sphe_ctx* c)
{
    if (!c->heights_stale || !c->device_valid) return SPHE_OK;
    sphe_launch_heights(c->D, c->stream);
    HIPCHK(c, hipGetLastError());
    c->heights_stale = false;
    return SPHE_OK;
}

int ensure_device(sphe_ctx* c)
{
    if (!c->layout_valid) {
        int r = build_layout(c);
        if (r) return r;
    }
    if (!c->device_valid) {
        int r = upload(c);
        if (r) return r;
    }
    return SPHE_OK;
}

int upload_table(sphe_ctx* c, bool force)
{
    std::vector<SceneDev> t(c->scenes.size());
    for (size_t i = 0; i < c->scenes.size(); ++i) {
        std::memset(&t[i], 0, sizeof(SceneDev));
        derive(c->scenes[i], t[i], c->yz);
    }
    if (!force && c->table_uploaded && t.size() == c->table.size() &&
        std::memcmp(t.data(), c->table.data(), t.size() * sizeof(SceneDev)) == 0)
        return SPHE_OK;
    c->table = t;
    if (!t.empty())
        HIPCHK(c, hipMemcpyAsync((void*)c->D.scenes, c->table.data(), t.size() * sizeof(SceneDev),
                                 hipMemcpyHostToDevice, c->stream));
    c->table_uploaded = true;
    ++c->epoch;  // new constants: a kept sort of the state no longer holds
    return SPHE_OK;
}

// The grid of a scene follows h and the box; a change of either after the
// layout was built needs a new layout.
int check_grid(sphe_ctx* c)
{
    for (Scene& s : c->scenes) {
        int dims[3];
        grid_dims(s.prm, dims, c->yz);
        if (dims[0] != s.dims[0] || dims[1] != s.dims[1] || dims[2] != s.dims[2]) {
            int r = begin_host_edit(c);
            if (r) return r;
            return ensure_device(c);
        }
    }
    return SPHE_OK;
}

int prepare(sphe_ctx* c)
{
    int r = ensure_device(c);
    if (r) return r;
    r = check_grid(c);
    if (r) return r;
    return upload_table(c, false);
}

// ---------------------------------------------------------------- steps

hipEvent_t get_event(sphe_ctx* c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Profiling: a pair of events the launcher hands to its first and last
// kernel (hipExtLaunchKernel), so they carry the kernels' own start and end
// times.  Without profiling both are null and launches are plain.
struct Timed {
    sphe_ctx* c;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    Timed(sphe_ctx* ctx, const char* n) : c(ctx), name(n)
    {
        if (c->profiling) {
            a = get_event(c);
            b = get_event(c);
        }
    }
    ~Timed()
    {
        if (c->profiling) c->pending.push_back(ProfRec{name, a, b});
    }
};

// block size of the force pass: its own when set, else the density pass's
int force_block(const sphe_ctx* c) { return c->fblock ? c->fblock : c->block; }

bool any_erosion(const sphe_ctx* c, int mode)
{
    for (const Scene& s : c->scenes)
        if (s.tnx > 0 && s.tnz > 0 && (mode < 0 ? s.prm.erosion != 0 : s.prm.erosion == mode)) return true;
    return false;
}

void launch_sort(sphe_ctx* c)
{
    const DevState& D = c->D;
    {
        Timed t(c, "scan");
        // with erosion the terrain heights of the last force pass are computed
        // in the scan's first launch (this step's force pass samples them)
        if (any_erosion(c, -1) && D.TC > 0) sphe_launch_scan_heights(D, c->stream, t.a, t.b);
        else sphe_launch_scan(D.cnt, D.start, D.C + 1, D.partial, false, D.ptop,
                              (unsigned long long)D.pool_origin, D.tocc, D.tfill, D.scan_log2, c->stream, t.a, t.b);
    }
    {
        Timed t(c, "scatter");
        sphe_launch_scatter(D, c->stream, t.a, t.b);
    }
    {
        Timed t(c, "reorder");
        sphe_launch_reorder(D, c->stream, t.a, t.b);
    }
}

void launch_step(sphe_ctx* c)
{
    const DevState& D = c->D;
    launch_sort(c);
    {
        Timed t(c, "density");
        sphe_launch_density(D, c->block, c->remap, c->lanes, c->stream, t.a, t.b);
    }
    {
        Timed t(c, "force");
        sphe_launch_force(D, force_block(c), 1, c->lanes, c->stream, t.a, t.b);  // plain XCD order: balance
    }
    if (any_erosion(c, SPHE_EROSION_TWO_PHASE)) {
        Timed t(c, "two_phase");
        sphe_launch_two_phase(D, c->stream, t.a, t.b);
    }
}

std::string plan_signature(const sphe_ctx* c)
{
    // everything the captured kernels take by value (DevState) or as launch shape
    char b[384];
    snprintf(b, sizeof(b), "%d/%d/%d/%d/%d/%lld/%lld/%lld/%d/%d/%p/%p/%p/%d/%d/%d/%d/%lld/%d/%d", c->precision, c->block,
             force_block(c), c->remap,
             c->lanes, (long long)c->D.P, (long long)c->D.C, (long long)c->D.TC, (int)any_erosion(c, -1),
             (int)any_erosion(c, SPHE_EROSION_TWO_PHASE), (const void*)c->D.u_pos, (const void*)c->D.nbl,
             (const void*)c->D.pool, c->D.maxn, c->D.sub_log2, c->D.yz, c->D.ocap, (long long)c->D.pool_cap,
             c->D.scan_log2, c->D.persistent);
    return b;
}

// The executable graph of a sequence of k steps (captured on first use; all
// sequences are dropped when the launch plan changes).  The sequence ends by
// publishing the list-overflow word to the host.
int ensure_graph(sphe_ctx* c, int k, hipGraphExec_t* exec)
{
    const std::string sig = plan_signature(c);
    if (sig != c->graph_sig) {
        drop_graphs(c);
        c->graph_sig = sig;
    }
    auto it = c->graphs.find(k);
    if (it != c->graphs.end()) {
        *exec = it->second.second;
        return SPHE_OK;
    }
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < k; ++i) launch_step(c);
    sphe_launch_publish(c->D, c->h_ovf_dev, c->stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess) return fail(c, SPHE_EHIP, "graph capture: %s", hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(c, SPHE_EHIP, "graph instantiate: %s", hipGetErrorString(e));
    }
    c->graphs[k] = std::make_pair(g, x);
    *exec = x;
    return SPHE_OK;
}

// Counts the cell of every unsorted slot from cleared counters.  After a
// re-upload the counters still hold the arrivals the last force pass made,
// so they must be cleared before the state is counted again.
int recount(sphe_ctx* c)
{
    HIPCHK(c, hipMemsetAsync(c->D.cnt, 0, (size_t)(c->D.C + 1) * 4, c->stream));
    {
        Timed t(c, "hash");
        sphe_launch_hash(c->D, c->stream, t.a, t.b);
    }
    HIPCHK(c, hipGetLastError());
    c->counted = true;
    return SPHE_OK;
}

// Replaces the neighbour lists (maxn entries per particle) and the overflow
// pool (chunks of ocap entries) where their sizes change.  The contents need
// not survive: every density pass rewrites them.  The old buffer is freed
// before the new one is taken, so the peak is the larger of the two; if the
// new size cannot be had, the old size is taken again and nothing changes
// (a particle without room then walks its stencil: the same result, slower).
int resize_lists(sphe_ctx* c, int maxn, int ocap, int64_t chunks)
{
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t P = c->D.P;
    const int64_t pool = pool_entries(c, chunks, ocap);
    auto replace = [&](void* old, size_t old_bytes, size_t new_bytes, void** out) -> bool {
        auto it = std::find(c->allocs.begin(), c->allocs.end(), old);
        (void)hipFree(old);
        void* p = nullptr;
        bool ok = true;
        if (try_malloc(&p, new_bytes) != hipSuccess) {
            ok = false;
            if (try_malloc(&p, old_bytes) != hipSuccess) p = nullptr;
        }
        if (it != c->allocs.end()) *it = p;
        *out = p;
        return ok;
    };
    // neither size to be had: the buffer is gone, so the particle state (in
    // other buffers, intact) goes to the host and the device layout is
    // dropped; the next call rebuilds it or reports the shortage
    auto lost = [&](const char* what) {
        (void)sync_to_host(c);
        free_device(c);
        return fail(c, SPHE_ENOMEM, "%s lost: no device memory (the state is kept on the host)", what);
    };
    bool ok = true;
    if (maxn != c->D.maxn) {
        void* p = nullptr;
        ok = replace(c->D.nbl, (size_t)P * c->D.maxn * 4, (size_t)P * maxn * 4, &p);
        c->D.nbl = (uint32_t*)p;
        if (!p) return lost("neighbour lists");
        if (ok) c->D.maxn = maxn;
    }
    if (ok && pool != c->D.pool_cap) {
        void* q = nullptr;
        ok = replace(c->D.pool, (size_t)c->D.pool_cap * 4, (size_t)pool * 4, &q);
        c->D.pool = (uint32_t*)q;
        if (!q) return lost("overflow pool");
        if (ok) c->D.pool_cap = pool;
    }
    if (ok) c->D.ocap = ocap;
    c->maxn = c->D.maxn;
    c->ocap = c->D.ocap;
    c->pool_chunks = (c->D.pool_cap - c->D.pool_origin) / c->D.ocap;
    ++c->epoch;  // the new lists hold nothing yet
    return SPHE_OK;
}

// Longest list a particle gets before its neighbours continue in a chunk of
// the overflow pool: the lists of all particles may take a quarter of the
// device's memory (a multiple of 128 entries, at least 256, at most 2^20).
// An overflowing particle costs two same-address atomics per step (the
// largest count and the chunk claim), so lists long enough for every
// neighbourhood are the fast case wherever memory allows them.
int list_max(const sphe_ctx* c)
{
    const int64_t per_entry = (c->D.P + 64) * 4;
    int64_t m = (int64_t)(c->mem_total / 4) / per_entry / 128 * 128;
    return (int)std::min<int64_t>(std::max<int64_t>(m, 256), 1 << 20);
}

// Grows the lists or the pool when the last step sequence published that a
// particle had no room: a count above the list and its chunk, or a chunk
// claim that found the pool full.  The published words are read without
// waiting for the stream; they lag by the launches still in flight, which
// only delays the growth.  Policy: lists up to 1.5x the largest count, at
// most list_max entries; longer neighbourhoods get chunks of 1.5x the excess
// over the list; the pool holds 1.5x the chunks the last step claimed.
int grow_lists(sphe_ctx* c)
{
    volatile uint32_t* h = c->h_ovf;
    const uint32_t seen = h[0], failed = h[1];
    if (seen <= (uint32_t)(c->D.maxn + c->D.ocap) && failed == 0) return SPHE_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    uint32_t w[2] = {0, 0};
    unsigned long long claimed = 0;
    HIPCHK(c, hipMemcpy(w, c->D.ovf, 8, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(&claimed, c->D.ptop, 8, hipMemcpyDeviceToHost));
    const int64_t need = std::max<int64_t>(w[0], seen);
    const int64_t fails = std::max<int64_t>(w[1], failed);
    auto up = [](int64_t x, int64_t m) { return (x + m - 1) / m * m; };
    int maxn = c->D.maxn;
    if (need > maxn) maxn = (int)std::max<int64_t>(maxn, std::min<int64_t>(up(need * 3 / 2, 128), list_max(c)));
    int ocap = c->D.ocap;
    if (need > maxn + ocap) ocap = (int)up((need - maxn) * 3 / 2, 128);
    const unsigned long long origin = (unsigned long long)c->D.pool_origin;
    const int64_t claims =
        (int64_t)((claimed > origin ? claimed - origin : 0ull) / (unsigned long long)c->D.ocap) + fails;
    int64_t chunks = (c->D.pool_cap - c->D.pool_origin) / c->D.ocap;
    if (fails > 0 || ocap != c->D.ocap) chunks = std::max<int64_t>(64, up(claims * 3 / 2, 64));
    chunks = std::min<int64_t>(chunks, (pool_max(c) - c->D.pool_origin) / ocap);
    const int r = resize_lists(c, maxn, ocap, chunks);
    if (r) return r;
    HIPCHK(c, hipMemsetAsync(c->D.ovf, 0, 8, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    clear_published(c);
    return SPHE_OK;
}

// Launches queued ahead of the one whose published overflow word the host
// reads (the size of sphe_ctx::seq_ev): the device never runs dry, and lists
// too short for the state grow at most this many launches late.
const int kInFlight = 2;

int do_steps(sphe_ctx* c, int nsteps)
{
    int r = prepare(c);
    if (r) return r;
    if (c->D.P == 0) return SPHE_OK;
    sphe_launch_clear_status(c->D, c->stream);
    if (!c->counted) {
        r = recount(c);
        if (r) return r;
    }
    HIPCHK(c, hipGetLastError());
    const bool graph = c->use_graph && !c->profiling;
    // Graph mode: sequences of c->seq_steps steps in one graph launch, the rest
    // one step per launch; direct launches (profiling, or graphs off) one
    // step at a time.  Before each launch the host looks at the overflow word
    // the last finished sequence published.
    for (int done = 0; done < nsteps;) {
        hipEvent_t& ev = c->seq_ev[c->seq_count % kInFlight];
        if (c->seq_count >= (uint64_t)kInFlight) HIPCHK(c, hipEventSynchronize(ev));
        r = grow_lists(c);
        if (r) return r;
        const int k = graph && nsteps - done >= c->seq_steps ? c->seq_steps : 1;
        if (graph) {
            hipGraphExec_t x = nullptr;
            r = ensure_graph(c, k, &x);
            if (r) return r;
            HIPCHK(c, hipGraphLaunch(x, c->stream));
        } else {
            launch_step(c);
            sphe_launch_publish(c->D, c->h_ovf_dev, c->stream);
        }
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipEventRecord(ev, c->stream));
        ++c->seq_count;
        done += k;
        if (any_erosion(c, -1)) c->heights_stale = true;
        c->stepped = true;
    }
    c->step_lanes = c->lanes;
    c->step_sub_log2 = c->D.sub_log2;
    c->idmap_valid = false;
    c->host_valid = false;
    ++c->epoch;
    return SPHE_OK;
}

int collect_profile(sphe_ctx* c)
{
    if (c->pending.empty()) return SPHE_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (ProfRec& p : c->pending) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, p.a, p.b));
        auto& e = c->prof[p.name];
        e.first += ms;
        e.second += 1;
        c->event_pool.push_back(p.a);
        c->event_pool.push_back(p.b);
    }
    c->pending.clear();
    return SPHE_OK;
}

// Sorts the current state and runs the density pass without stepping
// (verification hooks).  Kept until the state epoch changes, so the hooks of
// several scenes share one pass.
int sort_current(sphe_ctx* c)
{
    int r = prepare(c);
    if (r) return r;
    if (c->D.P == 0) return SPHE_OK;
    if (c->sorted_epoch == c->epoch && !c->stepped) return SPHE_OK;
    if (!c->counted) {
        r = recount(c);
        if (r) return r;
    }
    launch_sort(c);
    sphe_launch_density(c->D, c->block, c->remap, c->lanes, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->counted = false;
    c->stepped = false;
    c->idmap_valid = false;
    c->sorted_epoch = c->epoch;
    // the sorted buffers now hold the current state; keep the unsorted
    // state authoritative for the next step (its counts were consumed)
    return SPHE_OK;
}

Scene* scene_of(sphe_ctx* c, int scene)
{
    if (!c || scene < 0 || scene >= (int)c->scenes.size()) return nullptr;
    return &c->scenes[(size_t)scene];
}

const Scene* scene_of(const sphe_ctx* c, int scene)
{
    if (!c || scene < 0 || scene >= (int)c->scenes.size()) return nullptr;
    return &c->scenes[(size_t)scene];
}

#define SCENE_OR_FAIL(var, ctx, idx)                                                   \
    Scene* var = scene_of((ctx), (idx));                                               \
    if (!var) return fail((ctx), SPHE_ERANGE, "scene %d out of range", (int)(idx))

// Fixed-point unit 2^-fx_exp of the scene's volumes: chosen so that twice the
// tallest column the box allows stays below 2^50 units.  That leaves a factor
// of 2^12 below the largest volume a cell may hold (2^62): sediment carried
// from many cells can be deposited into one (a corner the water piles into),
// so a cell's volume is not bounded by its own column.
int set_terrain_heights(sphe_ctx* c, Scene& s, const float* heights)
{
    const size_t nt = (size_t)s.tnx * s.tnz;
    double top = (double)s.prm.box_center[1] + (double)s.prm.box_half;
    if (heights)
        for (size_t t = 0; t < nt; ++t) top = std::max(top, (double)heights[t]);
    const double col = std::max(1e-30, top - (double)s.t_org[1] + (double)s.t_cell);
    const double vmax = col * 2.0 * (double)s.t_cell * (double)s.t_cell;
    s.fx_exp = std::max(-60, std::min(100, 50 - (int)std::ceil(std::log2(vmax))));
    s.vol.assign(nt, 0);
    const double area = (double)s.t_cell * (double)s.t_cell;
    const double up = std::ldexp(1.0, s.fx_exp);
    for (size_t t = 0; t < nt; ++t) {
        const double hgt = heights ? (double)heights[t] : (double)s.t_org[1];
        if (!std::isfinite(hgt)) return fail(c, SPHE_EARG, "non-finite terrain height at cell %zu", t);
        const double v = std::max(0.0, hgt - (double)s.t_org[1]) * area * up;
        s.vol[t] = (int64_t)std::llround(v);
    }
    return SPHE_OK;
}

}  // namespace

// ====================================================================== ABI

extern "C" {

const char* sphe_version(void) { return "sphe 0.1.0 (gfx950)"; }

void sphe_params_default(sphe_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    // Erosion/fluid_system.h:460-478
    p->gravity[0] = 0.0f;
    p->gravity[1] = -9.82f;
    p->gravity[2] = 0.0f;
    p->dt = 0.0f;
    p->rest_density = 998.29f;
    p->mass = 0.02f;
    p->visc = 3.5f;
    p->surf_tens = 0.0728f;
    p->stiffness = 3.0f;
    p->h = 0.0457f;
    p->box_half = 0.2f;
    p->erosion = SPHE_EROSION_OFF;
    p->kc = 0.05f;
    p->ke = 0.3f;
    p->kd = 0.3f;
    p->band = 0.0457f;
}

int sphe_create(int precision, sphe_ctx** out)
{
    if (!out) return SPHE_EARG;
    *out = nullptr;
    if (precision < SPHE_FP32 || precision > SPHE_FP32_F64ACC) return SPHE_EARG;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return SPHE_EHIP;
    sphe_ctx* c = new sphe_ctx();
    c->precision = precision;
    c->device = dev;
    size_t mem_free = 0;
    if (hipMemGetInfo(&mem_free, &c->mem_total) != hipSuccess) {
        (void)hipGetLastError();
        c->mem_total = 0;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return SPHE_EHIP;
    }
    void* h = nullptr;
    if (hipHostMalloc(&h, 64, hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void**)&c->h_ovf_dev, h, 0) != hipSuccess) {
        (void)hipGetLastError();
        if (h) (void)hipHostFree(h);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return SPHE_EHIP;
    }
    c->h_ovf = (uint32_t*)h;
    clear_published(c);  // pinned memory may come back holding an earlier owner's words
    for (hipEvent_t& e : c->seq_ev) {
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            for (hipEvent_t& f : c->seq_ev)
                if (f) (void)hipEventDestroy(f);
            (void)hipHostFree(h);
            (void)hipStreamDestroy(c->stream);
            delete c;
            return SPHE_EHIP;
        }
    }
    // the static scratch's address on this (current) device; a failed lookup
    // leaves only sphe_mass_all_noalloc unavailable, which then reports it
    if (sphe_noalloc_scratch(dev, &c->noalloc_scratch, &c->noalloc_doubles) != 0) c->noalloc_scratch = nullptr;
    *out = c;
    return SPHE_OK;
}

int sphe_destroy(sphe_ctx* c)
{
    if (!c) return SPHE_EARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (ProfRec& p : c->pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    free_device(c);
    if (c->d_mesh) (void)hipFree(c->d_mesh);
    if (c->d_feed) (void)hipFree(c->d_feed);
    if (c->d_mass_all) (void)hipFree(c->d_mass_all);
    for (hipEvent_t e : c->seq_ev)
        if (e) (void)hipEventDestroy(e);
    if (c->h_ovf) (void)hipHostFree(c->h_ovf);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SPHE_OK;
}

const char* sphe_last_error(const sphe_ctx* c) { return c ? c->err.c_str() : "null context"; }

int sphe_add_scene(sphe_ctx* c, const sphe_params* p, const sphe_terrain_desc* t, int* scene)
{
    if (!c || !p) return fail(c, SPHE_EARG, "null argument");
    if (!(p->h > 0.0f) || !(p->box_half > 0.0f)) return fail(c, SPHE_EARG, "h and box_half must be > 0");
    int dims[3];
    grid_dims(*p, dims, c->yz);
    if ((double)dims[0] * dims[1] * dims[2] > 4.0e8) return fail(c, SPHE_ERANGE, "grid too fine for the box");
    int r = begin_host_edit(c);
    if (r) return r;
    Scene s;
    s.prm = *p;
    if (t && t->nx > 0 && t->nz > 0) {
        if (!(t->cell > 0.0f)) return fail(c, SPHE_EARG, "terrain cell must be > 0");
        s.tnx = t->nx;
        s.tnz = t->nz;
        for (int a = 0; a < 3; ++a) s.t_org[a] = t->origin[a];
        s.t_cell = t->cell;
        r = set_terrain_heights(c, s, t->heights);
        if (r) return r;
    }
    c->scenes.push_back(std::move(s));
    if (scene) *scene = (int)c->scenes.size() - 1;
    return SPHE_OK;
}

int sphe_scene_count(const sphe_ctx* c, int* count)
{
    if (!c || !count) return SPHE_EARG;
    *count = (int)c->scenes.size();
    return SPHE_OK;
}

int sphe_set_params(sphe_ctx* c, int scene, const sphe_params* p)
{
    if (!c || !p) return fail(c, SPHE_EARG, "null argument");
    SCENE_OR_FAIL(s, c, scene);
    if (!(p->h > 0.0f) || !(p->box_half > 0.0f)) return fail(c, SPHE_EARG, "h and box_half must be > 0");
    s->prm = *p;
    return SPHE_OK;
}

int sphe_get_params(const sphe_ctx* c, int scene, sphe_params* p)
{
    const Scene* s = scene_of(c, scene);
    if (!s || !p) return SPHE_EARG;
    *p = s->prm;
    return SPHE_OK;
}

float* sphe_param_ptr(sphe_ctx* c, int scene, int which)
{
    Scene* s = scene_of(c, scene);
    if (!s) return nullptr;
    switch (which) {
    case SPHE_P_MASS: return &s->prm.mass;
    case SPHE_P_VISC: return &s->prm.visc;
    case SPHE_P_SURF_TEN: return &s->prm.surf_tens;
    case SPHE_P_P0: return &s->prm.rest_density;
    case SPHE_P_GRAV: return s->prm.gravity;
    default: return nullptr;
    }
}

int sphe_set_dt(sphe_ctx* c, int scene, float dt)
{
    SCENE_OR_FAIL(s, c, scene);
    s->prm.dt = dt;
    return SPHE_OK;
}

int sphe_get_dt(const sphe_ctx* c, int scene, float* dt)
{
    const Scene* s = scene_of(c, scene);
    if (!s || !dt) return SPHE_EARG;
    *dt = s->prm.dt;
    return SPHE_OK;
}

int sphe_set_origin(sphe_ctx* c, int scene, const float o[3])
{
    SCENE_OR_FAIL(s, c, scene);
    if (!o) return fail(c, SPHE_EARG, "null origin");
    for (int a = 0; a < 3; ++a) s->origin[a] = o[a];
    return SPHE_OK;
}

int sphe_get_origin(const sphe_ctx* c, int scene, float o[3])
{
    const Scene* s = scene_of(c, scene);
    if (!s || !o) return SPHE_EARG;
    for (int a = 0; a < 3; ++a) o[a] = s->origin[a];
    return SPHE_OK;
}

int sphe_init_lattice(sphe_ctx* c, int scene, int n)
{
    SCENE_OR_FAIL(s, c, scene);
    if (n < 0) return fail(c, SPHE_EARG, "negative particle count");
    int r = begin_host_edit(c);
    if (r) return r;
    // Initialize (fluid_system.h:74-102): the id counter restarts at 0 and
    // the lattice is appended to whatever the scene holds.
    s->init_num = n;
    s->next_id = 0;
    append_lattice(*s, n);
    sort_by_id(*s);
    return SPHE_OK;
}

int sphe_add_particles(sphe_ctx* c, int scene, int n)
{
    SCENE_OR_FAIL(s, c, scene);
    if (n < 0) return fail(c, SPHE_EARG, "negative particle count");
    int r = begin_host_edit(c);
    if (r) return r;
    append_lattice(*s, n);
    sort_by_id(*s);
    return SPHE_OK;
}

int sphe_reset(sphe_ctx* c, int scene)
{
    SCENE_OR_FAIL(s, c, scene);
    int r = begin_host_edit(c);
    if (r) return r;
    // Reset (fluid_system.h:253-259)
    s->pos.clear();
    s->vel.clear();
    s->ids.clear();
    s->sed.clear();
    s->next_id = 0;
    append_lattice(*s, s->init_num);
    return SPHE_OK;
}

static int set_particles_any(sphe_ctx* c, int scene, int n, const float* pf, const double* pd, const float* vf,
                             const double* vd, const int32_t* ids, const int64_t* sed)
{
    SCENE_OR_FAIL(s, c, scene);
    if (n < 0 || (n > 0 && !pf && !pd)) return fail(c, SPHE_EARG, "bad particle arrays");
    if (ids)
        for (int i = 0; i < n; ++i)
            if (ids[i] < 0) return fail(c, SPHE_EARG, "negative particle id at %d", i);
    if (sed)
        for (int i = 0; i < n; ++i)
            if (sed[i] < 0) return fail(c, SPHE_EARG, "negative sediment at %d", i);
    int r = begin_host_edit(c);
    if (r) return r;
    s->pos.resize((size_t)3 * n);
    s->vel.assign((size_t)3 * n, 0.0);
    s->ids.resize((size_t)n);
    s->sed.assign((size_t)n, 0);
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            s->pos[3 * i + a] = pf ? (double)pf[3 * i + a] : pd[3 * i + a];
            if (vf) s->vel[3 * i + a] = (double)vf[3 * i + a];
            if (vd) s->vel[3 * i + a] = vd[3 * i + a];
        }
        s->ids[i] = ids ? ids[i] : i;
        if (sed) s->sed[i] = sed[i];
    }
    int32_t top = -1;
    for (int i = 0; i < n; ++i) top = std::max(top, s->ids[i]);
    s->next_id = top + 1;
    sort_by_id(*s);
    return SPHE_OK;
}

int sphe_set_particles(sphe_ctx* c, int scene, int n, const float* pos3, const float* vel3, const int32_t* ids,
                       const int64_t* sed_fx)
{
    return set_particles_any(c, scene, n, pos3, nullptr, vel3, nullptr, ids, sed_fx);
}

int sphe_set_particles_f64(sphe_ctx* c, int scene, int n, const double* pos3, const double* vel3,
                           const int32_t* ids, const int64_t* sed_fx)
{
    return set_particles_any(c, scene, n, nullptr, pos3, nullptr, vel3, ids, sed_fx);
}

int sphe_num_particles(const sphe_ctx* c, int scene, int* n)
{
    const Scene* s = scene_of(c, scene);
    if (!s || !n) return SPHE_EARG;
    *n = (int)s->ids.size();
    return SPHE_OK;
}

// density / pressure of the last step, in id order
static int read_rho(sphe_ctx* c, Scene& s, std::vector<double>& rho)
{
    const size_t n = s.ids.size();
    rho.assign(n, 0.0);
    if (!c->stepped || n == 0) return SPHE_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<int32_t> id(n);
    HIPCHK(c, hipMemcpy(id.data(), c->D.s_id + s.p_begin, n * 4, hipMemcpyDeviceToHost));
    std::vector<double> r(n);
    if (c->precision == SPHE_FP64) {
        HIPCHK(c, hipMemcpy(r.data(), (double*)c->D.rho + s.p_begin, n * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> rf(n);
        HIPCHK(c, hipMemcpy(rf.data(), (float*)c->D.rho + s.p_begin, n * 4, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) r[i] = rf[i];
    }
    std::vector<size_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return id[a] < id[b]; });
    for (size_t t = 0; t < n; ++t) rho[t] = r[perm[t]];
    return SPHE_OK;
}

enum { FEED_POS = 0, FEED_VEL = 1, FEED_RHO = 2 };

// Positions or velocities (3 floats per particle) or the last step's
// densities (1 float) of a scene whose ids are 0 .. n-1, in id order,
// gathered on the device from the current state into c->d_feed.
static int feed_on_device(sphe_ctx* c, const Scene* s, int what, float** out)
{
    const size_t nf = (what == FEED_RHO ? 1 : 3) * s->ids.size();
    int r = ensure_device(c);
    if (r) return r;
    if (nf > c->feed_cap) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->d_feed) (void)hipFree(c->d_feed);
        c->d_feed = nullptr;
        c->feed_cap = 0;
        HIPCHK(c, try_malloc(&c->d_feed, nf * 4));
        c->feed_cap = nf;
    }
    sphe_launch_feed(c->D, s->p_begin, (int64_t)s->ids.size(), what, c->d_feed, c->stream);
    HIPCHK(c, hipGetLastError());
    *out = c->d_feed;
    return SPHE_OK;
}

int sphe_read_buffer(sphe_ctx* c, int scene, int which, void* out, size_t bytes)
{
    SCENE_OR_FAIL(s, c, scene);
    const size_t n = s->ids.size();
    const size_t comps = (which == SPHE_BUF_POS || which == SPHE_BUF_VEL) ? 3 : 1;
    if (!out || bytes < n * comps * 4) return fail(c, SPHE_EARG, "output buffer too small");
    if (which < SPHE_BUF_POS || which > SPHE_BUF_ID) return fail(c, SPHE_EARG, "unknown buffer id %d", which);
    const bool state = which == SPHE_BUF_POS || which == SPHE_BUF_VEL;
    const bool rho = which == SPHE_BUF_DENSITY || which == SPHE_BUF_PRESS;
    float* f = (float*)out;
    if (n > 0 && s->ids_dense && c->device_valid && ((state && !c->host_valid) || (rho && c->stepped))) {
        // the device holds the values: gather this scene alone in id order
        // and copy 4 bytes per value, with no host sync of the rest
        float* d = nullptr;
        int r = feed_on_device(c, s, which == SPHE_BUF_POS ? FEED_POS : (state ? FEED_VEL : FEED_RHO), &d);
        if (r) return r;
        HIPCHK(c, hipMemcpyAsync(out, d, n * comps * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (which == SPHE_BUF_PRESS)
            for (size_t i = 0; i < n; ++i) f[i] = s->prm.stiffness * (f[i] - s->prm.rest_density);
        return SPHE_OK;
    }
    // the ids (ascending), the mass and the densities (from the device) need
    // no host copy of the state
    int r = state || which == SPHE_BUF_SED ? sync_to_host(c) : SPHE_OK;
    if (r) return r;
    switch (which) {
    case SPHE_BUF_POS:
        for (size_t i = 0; i < 3 * n; ++i) f[i] = (float)s->pos[i];
        break;
    case SPHE_BUF_VEL:
        for (size_t i = 0; i < 3 * n; ++i) f[i] = (float)s->vel[i];
        break;
    case SPHE_BUF_MASS:
        for (size_t i = 0; i < n; ++i) f[i] = s->prm.mass;
        break;
    case SPHE_BUF_SED: {
        const double unit = std::ldexp(1.0, -s->fx_exp);
        for (size_t i = 0; i < n; ++i) f[i] = (float)((double)s->sed[i] * unit);
        break;
    }
    case SPHE_BUF_ID:
        std::memcpy(out, s->ids.data(), n * 4);
        break;
    case SPHE_BUF_DENSITY:
    case SPHE_BUF_PRESS: {
        std::vector<double> rho;
        r = read_rho(c, *s, rho);
        if (r) return r;
        for (size_t i = 0; i < n; ++i)
            f[i] = which == SPHE_BUF_DENSITY ? (float)rho[i]
                                             : s->prm.stiffness * ((float)rho[i] - s->prm.rest_density);
        break;
    }
    default:
        return fail(c, SPHE_EARG, "unknown buffer id %d", which);
    }
    return SPHE_OK;
}

int sphe_buffer(sphe_ctx* c, int scene, int which, const void** view, size_t* count)
{
    SCENE_OR_FAIL(s, c, scene);
    if (!view) return fail(c, SPHE_EARG, "null view");
    const size_t n = s->ids.size();
    const size_t comps = (which == SPHE_BUF_POS || which == SPHE_BUF_VEL) ? 3 : 1;
    std::vector<float>& v = c->views[std::make_pair(scene, which)];
    v.assign(std::max<size_t>(1, n * comps), 0.0f);  // float and int32 records are 4 bytes
    int r = sphe_read_buffer(c, scene, which, v.data(), v.size() * 4);
    if (r) return r;
    *view = v.data();
    if (count) *count = n;
    return SPHE_OK;
}

int sphe_read_state_f64(sphe_ctx* c, int scene, double* pos3, double* vel3)
{
    SCENE_OR_FAIL(s, c, scene);
    int r = sync_to_host(c);
    if (r) return r;
    if (pos3) std::memcpy(pos3, s->pos.data(), s->pos.size() * 8);
    if (vel3) std::memcpy(vel3, s->vel.data(), s->vel.size() * 8);
    return SPHE_OK;
}

int sphe_read_sediment_fx(sphe_ctx* c, int scene, int64_t* sed)
{
    SCENE_OR_FAIL(s, c, scene);
    int r = sync_to_host(c);
    if (r) return r;
    if (!sed) return fail(c, SPHE_EARG, "null output");
    std::memcpy(sed, s->sed.data(), s->sed.size() * 8);
    return SPHE_OK;
}

int sphe_get_particle(sphe_ctx* c, int scene, int id, sphe_particle* out)
{
    SCENE_OR_FAIL(s, c, scene);
    if (!out) return fail(c, SPHE_EARG, "null output");
    std::memset(out, 0, sizeof(*out));
    out->Id = id;
    out->NeighbId = -1;
    if (!c->stepped || !c->device_valid) {
        // no step since the last edit: the staging arrays hold the state, and
        // there are no forces yet
        int r = sync_to_host(c);
        if (r) return r;
        const auto it = std::lower_bound(s->ids.begin(), s->ids.end(), id);
        if (it == s->ids.end() || *it != id) return fail(c, SPHE_ERANGE, "no particle with id %d", id);
        const size_t hi = (size_t)(it - s->ids.begin());
        for (int a = 0; a < 3; ++a) {
            out->Position[a] = (float)s->pos[3 * hi + a];
            out->Velocity[a] = (float)s->vel[3 * hi + a];
        }
        return SPHE_OK;
    }
    if (id < 0) return fail(c, SPHE_ERANGE, "no particle with id %d", id);
    // the slot of the id in the last step's sorted buffers: from the id map
    // (rebuilt once per step, on the first inspection) or by a search
    const int32_t* slotp = nullptr;
    if (s->idmap_off >= 0) {
        if (id >= s->idmap_len) return fail(c, SPHE_ERANGE, "no particle with id %d", id);
        if (!c->idmap_valid) {
            sphe_launch_idmap(c->D, c->d_idmap, c->idmap_len, c->stream);
            HIPCHK(c, hipGetLastError());
            c->idmap_valid = true;
        }
        slotp = c->d_idmap + s->idmap_off + id;
    } else {
        sphe_launch_find_id(c->D, scene, id, c->d_slot, c->stream);
        HIPCHK(c, hipGetLastError());
        slotp = c->d_slot;
    }
    // the sums in the order the step took them
    DevState D = c->D;
    D.sub_log2 = c->step_sub_log2;
    sphe_launch_inspect(D, slotp, c->step_lanes, c->d_inspect, c->stream);
    HIPCHK(c, hipGetLastError());
    double v[29];
    HIPCHK(c, hipMemcpyAsync(v, c->d_inspect, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (v[22] < 0.0) return fail(c, SPHE_ERANGE, "no particle with id %d", id);
    for (int a = 0; a < 3; ++a) {
        out->Position[a] = (float)v[23 + a];
        out->Velocity[a] = (float)v[26 + a];
    }
    out->Density = (float)v[0];
    out->Pressure = (float)v[1];
    for (int a = 0; a < 3; ++a) {
        out->Acceleration[a] = (float)v[2 + a];
        out->PressureForce[a] = (float)v[5 + a];
        out->ViscosityForce[a] = (float)v[8 + a];
        out->GravityForce[a] = (float)v[11 + a];
        out->SurfaceForce[a] = (float)v[14 + a];
        out->SurfaceNormal[a] = (float)v[17 + a];
    }
    out->NeighbId = (int32_t)v[20];
    return SPHE_OK;
}

int sphe_set_terrain(sphe_ctx* c, int scene, const sphe_terrain_desc* t)
{
    SCENE_OR_FAIL(s, c, scene);
    if (t && t->nx > 0 && t->nz > 0 && !(t->cell > 0.0f)) return fail(c, SPHE_EARG, "terrain cell must be > 0");
    int r = begin_host_edit(c);
    if (r) return r;
    if (!t || t->nx <= 0 || t->nz <= 0) {
        s->tnx = s->tnz = 0;
        s->vol.clear();
        return SPHE_OK;
    }
    s->tnx = t->nx;
    s->tnz = t->nz;
    for (int a = 0; a < 3; ++a) s->t_org[a] = t->origin[a];
    s->t_cell = t->cell;
    return set_terrain_heights(c, *s, t->heights);
}

int sphe_terrain_load_u8(sphe_ctx* c, int scene, const uint8_t* img, int w, int h, float scale)
{
    SCENE_OR_FAIL(s, c, scene);
    if (!img || w <= 0 || h <= 0) return fail(c, SPHE_EARG, "bad image");
    int r = begin_host_edit(c);
    if (r) return r;
    // The image is row-major, w pixels per row and h rows, as stbi_load
    // returns it (main.cpp:88-92), and Grid copies it unchanged
    // (grid.h:98-102).  Grid::GetHeightfieldAt(x, y) = map[dimY * x + y]
    // (grid.h:104-107) with dimY the row length: x picks the image row and y
    // the column.  So terrain cell (x, z) is pixel (row x, column z), and the
    // terrain has h cells along x and w along z.
    s->tnx = h;
    s->tnz = w;
    if (!(s->t_cell > 0.0f)) s->t_cell = 1.0f;
    std::vector<float> hts((size_t)w * h);
    for (int z = 0; z < w; ++z)
        for (int x = 0; x < h; ++x) hts[(size_t)z * h + x] = s->t_org[1] + scale * (float)img[(size_t)w * x + z];
    return set_terrain_heights(c, *s, hts.data());
}

int sphe_terrain_dims(const sphe_ctx* c, int scene, int* nx, int* nz)
{
    const Scene* s = scene_of(c, scene);
    if (!s) return SPHE_EARG;
    if (nx) *nx = s->tnx;
    if (nz) *nz = s->tnz;
    return SPHE_OK;
}

int sphe_fixed_exponent(const sphe_ctx* c, int scene, int* e)
{
    const Scene* s = scene_of(c, scene);
    if (!s || !e) return SPHE_EARG;
    *e = s->fx_exp;
    return SPHE_OK;
}

int sphe_terrain_volume_fx(sphe_ctx* c, int scene, int64_t* out, size_t count)
{
    SCENE_OR_FAIL(s, c, scene);
    int r = sync_to_host(c);
    if (r) return r;
    if (!out || count < s->vol.size()) return fail(c, SPHE_EARG, "output too small");
    std::memcpy(out, s->vol.data(), s->vol.size() * 8);
    return SPHE_OK;
}

int sphe_set_terrain_volume_fx(sphe_ctx* c, int scene, const int64_t* vol, size_t count)
{
    SCENE_OR_FAIL(s, c, scene);
    const size_t nt = (size_t)s->tnx * s->tnz;
    if (!vol || nt == 0 || count != nt) return fail(c, SPHE_EARG, "expected %zu cell volumes", nt);
    // erosion is not capped at a cell's volume, so a cell can go below the
    // terrain base (a negative volume): any value in the fixed-point range
    for (size_t t = 0; t < nt; ++t)
        if (vol[t] < -((int64_t)1 << 62) || vol[t] > ((int64_t)1 << 62))
            return fail(c, SPHE_ERANGE, "cell %zu volume out of range", t);
    int r = begin_host_edit(c);
    if (r) return r;
    std::memcpy(s->vol.data(), vol, nt * 8);
    return SPHE_OK;
}

int sphe_terrain_heights(sphe_ctx* c, int scene, float* out, size_t count)
{
    SCENE_OR_FAIL(s, c, scene);
    const size_t nt = (size_t)s->tnx * s->tnz;
    if (!out || count < nt) return fail(c, SPHE_EARG, "output too small");
    if (nt == 0) return SPHE_OK;
    int r = ensure_device(c);
    if (!r) r = ensure_heights(c);
    if (r) return r;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(out, c->D.height + s->t_off, nt * 4, hipMemcpyDeviceToHost));
    return SPHE_OK;
}

// The scene's terrain vertex stream computed on the device (k_terrain_mesh)
// into the context's mesh buffer, which grows as needed.
int mesh_on_device(sphe_ctx* c, const Scene* s, float** out)
{
    const size_t nv = (size_t)s->tnx * s->tnz * 6;
    int r = ensure_device(c);
    if (!r) r = ensure_heights(c);
    if (r) return r;
    if (nv > c->mesh_cap) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->d_mesh) (void)hipFree(c->d_mesh);
        c->d_mesh = nullptr;
        c->mesh_cap = 0;
        HIPCHK(c, try_malloc(&c->d_mesh, nv * 4));
        c->mesh_cap = nv;
    }
    sphe_launch_terrain_mesh(c->D.height + s->t_off, s->tnx, s->tnz, s->t_cell, s->t_org[0], s->t_org[2],
                             c->d_mesh, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = c->d_mesh;
    return SPHE_OK;
}

int sphe_terrain_mesh(sphe_ctx* c, int scene, float* verts, size_t* nverts, uint32_t* indices, size_t* nindices)
{
    SCENE_OR_FAIL(s, c, scene);
    const int nx = s->tnx, nz = s->tnz;
    const size_t nv = (size_t)nx * nz * 6;
    // genIndices (grid.h:118-136): two triangles per quad, the second walked
    // from the far corner
    size_t ni = 0;
    for (int z = 0, j = nz - 1; z < nz && j >= 0; ++z, --j)
        for (int x = 0, i = nx - 1; x < nx && i >= 0; ++x, --i) {
            if (x + 1 < nx && z + 1 < nz) ni += 3;
            if (j - 1 >= 0 && i - 1 >= 0) ni += 3;
        }
    if (nverts && !verts) *nverts = nv;
    if (nindices && !indices) *nindices = ni;
    if (!verts && !indices) return SPHE_OK;
    if ((verts && (!nverts || *nverts < nv)) || (indices && (!nindices || *nindices < ni)))
        return fail(c, SPHE_EARG, "mesh output too small");
    if (verts) {
        // UpdateGrid vertex stream (grid.h:138-176), from the device
        if (nv > 0) {
            float* d = nullptr;
            int r = mesh_on_device(c, s, &d);
            if (r) return r;
            HIPCHK(c, hipMemcpy(verts, d, nv * 4, hipMemcpyDeviceToHost));
        }
        *nverts = nv;
    }
    if (indices) {
        size_t k = 0;
        for (int z = 0, j = nz - 1; z < nz && j >= 0; ++z, --j)
            for (int x = 0, i = nx - 1; x < nx && i >= 0; ++x, --i) {
                if (x + 1 < nx && z + 1 < nz) {
                    indices[k++] = (uint32_t)(z * nx + x);
                    indices[k++] = (uint32_t)(z * nx + x + 1);
                    indices[k++] = (uint32_t)((z + 1) * nx + x);
                }
                if (j - 1 >= 0 && i - 1 >= 0) {
                    indices[k++] = (uint32_t)(j * nx + i);
                    indices[k++] = (uint32_t)(j * nx + i - 1);
                    indices[k++] = (uint32_t)((j - 1) * nx + i);
                }
            }
        *nindices = ni;
    }
    return SPHE_OK;
}

int sphe_terrain_mesh_device(sphe_ctx* c, int scene, const float** dverts, size_t* nfloats)
{
    SCENE_OR_FAIL(s, c, scene);
    if (!dverts || !nfloats) return fail(c, SPHE_EARG, "null output");
    const size_t nv = (size_t)s->tnx * s->tnz * 6;
    *nfloats = nv;
    *dverts = nullptr;
    if (nv == 0) return SPHE_OK;
    float* d = nullptr;
    int r = mesh_on_device(c, s, &d);
    if (r) return r;
    *dverts = d;
    return SPHE_OK;
}

int sphe_positions_device(sphe_ctx* c, int scene, const float** dpos, size_t* nfloats)
{
    SCENE_OR_FAIL(s, c, scene);
    if (!dpos || !nfloats) return fail(c, SPHE_EARG, "null output");
    *dpos = nullptr;
    *nfloats = 3 * s->ids.size();
    if (s->ids.empty()) return SPHE_OK;
    if (!c->layout_valid) {
        int r = build_layout(c);  // decides whether the ids are 0 .. n-1
        if (r) return r;
    }
    if (!s->ids_dense) return fail(c, SPHE_EARG, "the scene's ids are not 0 .. n-1");
    float* d = nullptr;
    int r = feed_on_device(c, s, FEED_POS, &d);
    if (r) return r;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *dpos = d;
    return SPHE_OK;
}

int sphe_batch_create(int precision, const sphe_params* params, const sphe_terrain_desc* terrains, int n,
                      sphe_ctx** out)
{
    if (!out || n < 0 || (n > 0 && !params)) return SPHE_EARG;
    *out = nullptr;
    sphe_ctx* c = nullptr;
    int r = sphe_create(precision, &c);
    if (r) return r;
    for (int i = 0; i < n; ++i) {
        int sid = 0;
        r = sphe_add_scene(c, &params[i], terrains ? &terrains[i] : nullptr, &sid);
        if (r) {
            (void)sphe_destroy(c);
            return r;
        }
    }
    *out = c;
    return SPHE_OK;
}

int sphe_batch_step(sphe_ctx* c, int nsteps) { return sphe_step(c, nsteps); }

int sphe_step(sphe_ctx* c, int nsteps)
{
    if (!c) return SPHE_EARG;
    if (nsteps < 0) return fail(c, SPHE_EARG, "negative step count");
    if (nsteps == 0) return SPHE_OK;
    return do_steps(c, nsteps);
}

int sphe_sync(sphe_ctx* c)
{
    if (!c) return SPHE_EARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return collect_profile(c);
}

int sphe_use_graph(sphe_ctx* c, int enable)
{
    if (!c) return SPHE_EARG;
    c->use_graph = enable != 0;
    return SPHE_OK;
}

int sphe_mass(sphe_ctx* c, int scene, double* terrain, double* sediment, uint64_t* units)
{
    SCENE_OR_FAIL(s, c, scene);
    int r = prepare(c);
    if (r) return r;
    sphe_launch_mass(c->D, scene, c->d_mass_part, c->d_mass_out, c->stream);
    HIPCHK(c, hipGetLastError());
    double v[3];
    HIPCHK(c, hipMemcpyAsync(v, c->d_mass_out, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (terrain) *terrain = v[0];
    if (sediment) *sediment = v[1];
    if (units) std::memcpy(units, &v[2], 8);
    (void)s;
    return SPHE_OK;
}

int sphe_mass_all(sphe_ctx* c, double* terrain, double* sediment, uint64_t* units, int count)
{
    if (!c) return SPHE_EARG;
    const int ns = (int)c->scenes.size();
    if (count < ns) return fail(c, SPHE_EARG, "outputs hold %d scenes, the context has %d", count, ns);
    if (ns > 65535) return fail(c, SPHE_ERANGE, "more than 65535 scenes");
    if (ns == 0) return SPHE_OK;
    int r = prepare(c);
    if (r) return r;
    const size_t part = (size_t)3 * SPHE_MASS_BLOCKS_PER_SCENE * ns, outn = (size_t)3 * ns;
    if (c->mass_all_cap < part + outn) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (c->d_mass_all) (void)hipFree(c->d_mass_all);
        c->d_mass_all = nullptr;
        c->mass_all_cap = 0;
        HIPCHK(c, try_malloc(&c->d_mass_all, (part + outn) * 8));
        c->mass_all_cap = part + outn;
    }
    double* out = c->d_mass_all + part;
    sphe_launch_mass_scenes(c->D, c->d_mass_all, out, c->stream);
    HIPCHK(c, hipGetLastError());
    // planar results: terrain volumes, sediment, unit sums of all scenes
    if (terrain) HIPCHK(c, hipMemcpyAsync(terrain, out, (size_t)ns * 8, hipMemcpyDeviceToHost, c->stream));
    if (sediment) HIPCHK(c, hipMemcpyAsync(sediment, out + ns, (size_t)ns * 8, hipMemcpyDeviceToHost, c->stream));
    if (units) HIPCHK(c, hipMemcpyAsync(units, out + 2 * ns, (size_t)ns * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SPHE_OK;
}

int sphe_mass_all_noalloc(sphe_ctx* c, double* terrain, double* sediment, uint64_t* units, int count)
{
    if (!c) return SPHE_EARG;
    const int ns = (int)c->scenes.size();
    if (count < ns) return fail(c, SPHE_EARG, "outputs hold %d scenes, the context has %d", count, ns);
    if (ns > SPHE_NOALLOC_MAX_SCENES)
        return fail(c, SPHE_ERANGE, "sphe_mass_all_noalloc serves at most %d scenes, the context has %d",
                    SPHE_NOALLOC_MAX_SCENES, ns);
    if (!c->noalloc_scratch) return fail(c, SPHE_EHIP, "the library's static scratch is not available on this device");
    if (ns == 0) return SPHE_OK;
    // the layout and the device state are not built here (that allocates)
    if (!c->layout_valid || !c->device_valid)
        return fail(c, SPHE_ESTATE, "no device state to reduce: step the context (or call sphe_mass_all) first");
    const size_t part = (size_t)3 * SPHE_MASS_BLOCKS_PER_SCENE * ns;
    if (part + (size_t)3 * ns > c->noalloc_doubles) return fail(c, SPHE_ERANGE, "static scratch too small");
    // every partial and result is rewritten by the launch: nothing to clear
    double* out = c->noalloc_scratch + part;
    sphe_launch_mass_scenes(c->D, c->noalloc_scratch, out, c->stream);
    HIPCHK(c, hipGetLastError());
    if (terrain) HIPCHK(c, hipMemcpyAsync(terrain, out, (size_t)ns * 8, hipMemcpyDeviceToHost, c->stream));
    if (sediment) HIPCHK(c, hipMemcpyAsync(sediment, out + ns, (size_t)ns * 8, hipMemcpyDeviceToHost, c->stream));
    if (units) HIPCHK(c, hipMemcpyAsync(units, out + 2 * ns, (size_t)ns * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SPHE_OK;
}

int sphe_status(sphe_ctx* c, uint32_t* nonfinite)
{
    if (!c || !nonfinite) return SPHE_EARG;
    *nonfinite = 0;
    if (!c->layout_valid) return SPHE_OK;
    HIPCHK(c, hipMemcpyAsync(nonfinite, c->D.status, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SPHE_OK;
}

int sphe_profile(sphe_ctx* c, int enable)
{
    if (!c) return SPHE_EARG;
    int r = collect_profile(c);
    if (r) return r;
    c->profiling = enable != 0;
    c->prof.clear();
    return SPHE_OK;
}

int sphe_profile_read(sphe_ctx* c, const char* kernel, double* avg_ms, int* launches)
{
    if (!c || !kernel) return SPHE_EARG;
    int r = collect_profile(c);
    if (r) return r;
    auto it = c->prof.find(kernel);
    if (it == c->prof.end()) {
        if (avg_ms) *avg_ms = 0.0;
        if (launches) *launches = 0;
        return SPHE_OK;
    }
    if (avg_ms) *avg_ms = it->second.second ? it->second.first / it->second.second : 0.0;
    if (launches) *launches = it->second.second;
    return SPHE_OK;
}

int sphe_profile_names(sphe_ctx* c, char* names, size_t cap)
{
    if (!c || !names || cap == 0) return SPHE_EARG;
    int r = collect_profile(c);
    if (r) return r;
    std::string s;
    for (auto& kv : c->prof) {
        if (!s.empty()) s += ",";
        s += kv.first;
    }
    std::snprintf(names, cap, "%s", s.c_str());
    return SPHE_OK;
}

int sphe_kernel_bytes(sphe_ctx* c, const char* kernel, double* bytes)
{
    if (!c || !kernel || !bytes) return SPHE_EARG;
    const double E = (double)real_size(c->precision);
    double N = 0, cells = 0, tcells = 0, terr_n = 0, ero_n = 0;
    for (const Scene& s : c->scenes) {
        const double n = (double)s.ids.size();
        N += n;
        int dims[3];
        grid_dims(s.prm, dims, c->yz);
        cells += (double)dims[0] * dims[1] * dims[2] + 1.0;
        tcells += (double)s.tnx * s.tnz;
        if (s.tnx > 0) terr_n += n;
        if (s.tnx > 0 && s.prm.erosion) ero_n += n;
    }
    const std::string k = kernel;
    // compulsory inputs and outputs of each pass (intermediates kept in
    // registers or on chip are not counted; the cell table is the grid's
    // own data, counted once in the scan, so no pass's figure depends on
    // how finely the grid is cut)
    if (k == "density") *bytes = N * (3 * E + E);                       // pos -> rho
    else if (k == "force")
        *bytes = N * (3 * E + 3 * E + E) + N * (3 * E + 3 * E) + N * 8  // pos, vel, rho -> pos, vel; key+rank
                 + terr_n * 4 + ero_n * (8 + 8);                        // heights; sediment in/out
    else if (k == "reorder") *bytes = N * (12 + 2 * (6 * E + 12));       // tmp + gather + write
    else if (k == "scatter") *bytes = N * (12 + 12 + 4);
    else if (k == "scan") *bytes = 12 * cells;
    else if (k == "heights") *bytes = 12 * tcells;
    else if (k == "hash") *bytes = N * (3 * E + 8);
    else return fail(c, SPHE_EARG, "unknown kernel %s", kernel);
    return SPHE_OK;
}

int sphe_set_tuning(sphe_ctx* c, int block, int remap)
{
    if (!c) return SPHE_EARG;
    if (block != 64 && block != 128 && block != 256 && block != 512)
        return fail(c, SPHE_EARG, "block must be 64, 128, 256 or 512");
    if (remap < 0) return fail(c, SPHE_EARG, "xcd_remap (blocks per XCD run) must be >= 0");
    c->block = block;
    c->fblock = 0;
    c->remap = remap;
    return SPHE_OK;
}

int sphe_set_force_block(sphe_ctx* c, int block)
{
    if (!c) return SPHE_EARG;
    if (block != 0 && block != 64 && block != 128 && block != 256 && block != 512)
        return fail(c, SPHE_EARG, "force block must be 0, 64, 128, 256 or 512");
    c->fblock = block;
    return SPHE_OK;
}

int sphe_get_force_block(const sphe_ctx* c, int* block)
{
    if (!c) return SPHE_EARG;
    if (block) *block = force_block(c);
    return SPHE_OK;
}

int sphe_set_lanes(sphe_ctx* c, int lanes)
{
    if (!c) return SPHE_EARG;
    if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8) return fail(c, SPHE_EARG, "lanes must be 1, 2, 4 or 8");
    if (lanes != c->lanes) ++c->epoch;  // a kept density pass used the old split
    c->lanes = lanes;
    return SPHE_OK;
}

int sphe_get_lanes(const sphe_ctx* c, int* lanes)
{
    if (!c || !lanes) return SPHE_EARG;
    *lanes = c->lanes;
    return SPHE_OK;
}

int sphe_set_list_capacity(sphe_ctx* c, int entries)
{
    if (!c) return SPHE_EARG;
    if (entries < 16 || entries > (1 << 20) || entries % 16)
        return fail(c, SPHE_EARG, "list capacity must be a multiple of 16 in [16, 1048576]");
    // the overflow chunks take the same size; their count is the initial one
    if (!c->layout_valid) {
        c->maxn = entries;
        c->ocap = entries;
        c->pool_chunks = 0;
        return SPHE_OK;
    }
    if (entries == c->D.maxn && entries == c->D.ocap) return SPHE_OK;
    const int r = resize_lists(c, entries, entries, pool_chunks_initial(c->D.P, entries));
    if (r) return r;
    if (c->D.maxn != entries || c->D.ocap != entries)
        return fail(c, SPHE_ENOMEM, "no device memory for lists of %d entries", entries);
    return SPHE_OK;
}

int sphe_get_list_capacity(const sphe_ctx* c, int* entries)
{
    if (!c || !entries) return SPHE_EARG;
    *entries = c->layout_valid ? c->D.maxn : c->maxn;
    return SPHE_OK;
}

int sphe_get_overflow_capacity(const sphe_ctx* c, int* chunk_entries, int64_t* chunks)
{
    if (!c) return SPHE_EARG;
    const int oc = c->layout_valid ? c->D.ocap : c->ocap;
    if (chunk_entries) *chunk_entries = oc;
    if (chunks) *chunks = c->layout_valid ? (c->D.pool_cap - c->D.pool_origin) / oc : 0;
    return SPHE_OK;
}

int sphe_set_sequence_steps(sphe_ctx* c, int steps)
{
    if (!c) return SPHE_EARG;
    if (steps < 1 || steps > 4096) return fail(c, SPHE_EARG, "steps per sequence must be in [1, 4096]");
    c->seq_steps = steps;  // graphs are kept per length: nothing else changes
    return SPHE_OK;
}

int sphe_get_sequence_steps(const sphe_ctx* c, int* steps)
{
    if (!c || !steps) return SPHE_EARG;
    *steps = c->seq_steps;
    return SPHE_OK;
}

int sphe_debug_pool_origin(sphe_ctx* c, int64_t entries)
{
    if (!c) return SPHE_EARG;
    if (entries < 0 || entries % SPHE_POOL_UNIT || entries > pool_max(c) / 2)
        return fail(c, SPHE_EARG, "pool origin must be a multiple of %d in [0, %lld]", SPHE_POOL_UNIT,
                    (long long)(pool_max(c) / 2));
    if (entries == c->pool_origin) return SPHE_OK;
    // a new pool: the state goes through the host and the layout is rebuilt
    int r = begin_host_edit(c);
    if (r) return r;
    c->pool_origin = entries;
    return SPHE_OK;
}

int sphe_set_subcells(sphe_ctx* c, int slabs)
{
    if (!c) return SPHE_EARG;
    if (slabs != 1 && slabs != 2 && slabs != 4) return fail(c, SPHE_EARG, "slabs must be 1, 2 or 4");
    int sl = slabs == 4 ? 2 : (slabs == 2 ? 1 : 0);
    c->sub_log2 = sl;
    if (c->layout_valid) {
        while (sl > 0 && (c->D.C << sl) > INT32_MAX) --sl;
        if (sl != c->D.sub_log2) {
            c->D.sub_log2 = sl;
            ++c->epoch;
            c->counted = false;  // the keys of the counted state carry the slab
        }
    }
    return SPHE_OK;
}

int sphe_get_subcells(const sphe_ctx* c, int* slabs)
{
    if (!c || !slabs) return SPHE_EARG;
    *slabs = 1 << (c->layout_valid ? c->D.sub_log2 : c->sub_log2);
    return SPHE_OK;
}

int sphe_set_yz_split(sphe_ctx* c, int split)
{
    if (!c) return SPHE_EARG;
    if (split < 1 || split > 3) return fail(c, SPHE_EARG, "yz split must be 1, 2 or 3");
    if (split == c->yz) return SPHE_OK;
    for (const Scene& s : c->scenes) {
        int dims[3];
        grid_dims(s.prm, dims, split);
        if ((double)dims[0] * dims[1] * dims[2] > 4.0e8) return fail(c, SPHE_ERANGE, "grid too fine for the box");
    }
    // a new cell grid: the state goes through the host and the layout is rebuilt
    int r = begin_host_edit(c);
    if (r) return r;
    c->yz = split;
    return SPHE_OK;
}

int sphe_get_yz_split(const sphe_ctx* c, int* split)
{
    if (!c || !split) return SPHE_EARG;
    *split = c->yz;
    return SPHE_OK;
}

int sphe_set_scan_tile(sphe_ctx* c, int log2_cells)
{
    if (!c) return SPHE_EARG;
    if (log2_cells < SPHE_SCAN_LOG2_MIN || log2_cells > SPHE_SCAN_LOG2_MAX)
        return fail(c, SPHE_EARG, "scan tile must be 2^%d .. 2^%d cells", SPHE_SCAN_LOG2_MIN, SPHE_SCAN_LOG2_MAX);
    c->scan_log2 = log2_cells;
    if (!c->layout_valid || c->D.scan_log2 == log2_cells) return SPHE_OK;
    // the tile flags and fill words describe tiles of the old size: every
    // tile is read and written by the next scan, which sets them up again
    HIPCHK(c, hipMemsetAsync(c->D.tocc, 0x01, (size_t)c->D.npartial * 4, c->stream));
    HIPCHK(c, hipMemsetAsync(c->D.tfill, 0xFF, (size_t)c->D.npartial * 4, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->D.scan_log2 = log2_cells;
    return SPHE_OK;
}

int sphe_set_persistent(sphe_ctx* c, int on)
{
    if (!c) return SPHE_EARG;
    c->persistent = on != 0;
    if (c->layout_valid) c->D.persistent = c->persistent ? 1 : 0;
    return SPHE_OK;
}

int sphe_get_persistent(const sphe_ctx* c, int* on)
{
    if (!c || !on) return SPHE_EARG;
    *on = c->persistent ? 1 : 0;
    return SPHE_OK;
}

int sphe_get_scan_tile(const sphe_ctx* c, int* log2_cells)
{
    if (!c || !log2_cells) return SPHE_EARG;
    *log2_cells = c->layout_valid ? c->D.scan_log2 : c->scan_log2;
    return SPHE_OK;
}

int sphe_get_tuning(const sphe_ctx* c, int* block, int* remap)
{
    if (!c) return SPHE_EARG;
    if (block) *block = c->block;
    if (remap) *remap = c->remap;
    return SPHE_OK;
}

int sphe_save_state(sphe_ctx* c)
{
    if (!c) return SPHE_EARG;
    int r = sync_to_host(c);
    if (r) return r;
    c->saved.clear();
    for (const Scene& s : c->scenes) c->saved.push_back({s.pos, s.vel, s.ids, s.sed, s.vol, s.next_id});
    c->has_saved = true;
    return SPHE_OK;
}

int sphe_restore_state(sphe_ctx* c)
{
    if (!c) return SPHE_EARG;
    if (!c->has_saved) return fail(c, SPHE_ESTATE, "no saved state");
    if (c->saved.size() != c->scenes.size()) return fail(c, SPHE_ESTATE, "scenes were added after the state was saved");
    bool same = true;
    for (size_t i = 0; i < c->scenes.size(); ++i)
        same = same && c->scenes[i].ids.size() == c->saved[i].ids.size() &&
               c->scenes[i].vol.size() == c->saved[i].vol.size();
    if (!same) {
        int r = begin_host_edit(c);  // particle or terrain sizes changed: new layout
        if (r) return r;
    }
    for (size_t i = 0; i < c->scenes.size(); ++i) {
        Scene& s = c->scenes[i];
        const auto& v = c->saved[i];
        s.pos = v.pos; s.vel = v.vel; s.ids = v.ids; s.sed = v.sed; s.vol = v.vol; s.next_id = v.next_id;
    }
    c->host_valid = true;
    c->device_valid = false;  // uploaded (and counted) on the next use
    ++c->epoch;
    c->stepped = false;
    return SPHE_OK;
}

int sphe_autotune(sphe_ctx* c, int steps_per_trial)
{
    if (!c) return SPHE_EARG;
    if (steps_per_trial < 1) steps_per_trial = 1;
    static std::mutex mu;
    struct Tuned {
        int block, fblock, remap;
    };
    static std::map<std::string, Tuned> cache;
    int r = prepare(c);
    if (r) return r;
    hipDeviceProp_t prop;
    HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
    // what the choice depends on: the device, the kernels' variants and the sizes (not the tuning)
    std::string key = std::string(prop.gcnArchName) + "|" + std::to_string(c->precision) + "/" +
                      std::to_string(c->lanes) + "/" + std::to_string(c->remap) + "/" + std::to_string(c->D.P) + "/" +
                      std::to_string(c->D.C) + "/" + std::to_string(c->D.maxn) + "/" + std::to_string(c->D.sub_log2) +
                      "/" + std::to_string(c->D.yz) + "/" + std::to_string((int)any_erosion(c, -1));
    for (const Scene& s : c->scenes) key += "|" + std::to_string(s.ids.size()) + ":" + std::to_string(s.prm.h);
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            c->block = it->second.block;
            c->fblock = it->second.fblock;
            c->remap = it->second.remap;
            return SPHE_OK;
        }
    }
    r = sync_to_host(c);
    if (r) return r;
    // each trial times the density and force passes with the profiling
    // events (direct launches), and each pass keeps its own fastest block
    r = collect_profile(c);
    if (r) return r;
    const bool prof = c->profiling;
    const auto saved_prof = c->prof;
    int best_b = c->block, best_fb = force_block(c);
    double best_d = 1e30, best_f = 1e30;
    // the block sizes only: the best XCD run length depends on how uneven the
    // state is (long runs win on uniform states and lose by up to a fifth on
    // pile-ups), which a trial on the current state cannot foresee; the run
    // stays as set (default 128, near the best on both kinds of state)
    for (int b : {64, 128, 256, 512}) {
        c->block = b;
        c->fblock = b;
        c->device_valid = false;  // every trial starts from the same state
        c->profiling = false;
        r = do_steps(c, 1);
        if (r == SPHE_OK && hipStreamSynchronize(c->stream) != hipSuccess) r = fail(c, SPHE_EHIP, "autotune sync");
        if (r == SPHE_OK) {
            c->prof.clear();
            c->profiling = true;
            r = do_steps(c, steps_per_trial);
            if (r == SPHE_OK) r = collect_profile(c);
            c->profiling = false;
        }
        if (r) {
            c->prof = saved_prof;
            c->profiling = prof;
            return r;
        }
        const auto d = c->prof.find("density"), f = c->prof.find("force");
        const double dm = d != c->prof.end() ? d->second.first : 0.0, fm = f != c->prof.end() ? f->second.first : 0.0;
        if (dm < best_d) { best_d = dm; best_b = b; }
        if (fm < best_f) { best_f = fm; best_fb = b; }
    }
    c->prof = saved_prof;
    c->block = best_b;
    c->fblock = best_fb;
    c->profiling = prof;
    c->host_valid = true;     // host staging still holds the pre-tuning state
    c->device_valid = false;  // restore it on the next use
    c->stepped = false;
    {
        std::lock_guard<std::mutex> g(mu);
        cache[key] = {c->block, c->fblock, c->remap};
    }
    return SPHE_OK;
}

// One trial of the grid autotuner: the step from the state in host staging
// (one untimed step that also grows the lists, then `steps` profiled
// steps); *ms = the kernels' time per step.
static int grid_trial(sphe_ctx* c, int steps, double* ms)
{
    c->host_valid = true;     // staging holds the state every trial starts from
    c->device_valid = false;
    c->profiling = false;
    int r = do_steps(c, 2);
    if (r == SPHE_OK && hipStreamSynchronize(c->stream) != hipSuccess) r = fail(c, SPHE_EHIP, "autotune sync");
    if (r) return r;
    c->host_valid = true;
    c->device_valid = false;
    c->prof.clear();
    c->profiling = true;
    r = do_steps(c, steps);
    if (r == SPHE_OK) r = collect_profile(c);
    c->profiling = false;
    if (r) return r;
    double t = 0.0;
    for (const auto& kv : c->prof) t += kv.second.first;
    *ms = t / steps;
    return SPHE_OK;
}

int sphe_autotune_grid(sphe_ctx* c, int steps_per_trial)
{
    if (!c) return SPHE_EARG;
    if (steps_per_trial < 1) steps_per_trial = 1;
    static std::mutex mu;
    struct Tuned {
        int yz, slabs, scan_log2;
    };
    static std::map<std::string, Tuned> cache;
    int r = prepare(c);
    if (r) return r;
    hipDeviceProp_t prop;
    HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
    // what the choice depends on: the device, the kernels' variants, the
    // scenes' sizes and smoothing lengths (not the grid being tuned)
    std::string key = std::string(prop.gcnArchName) + "|" + std::to_string(c->precision) + "/" +
                      std::to_string(c->lanes) + "/" + std::to_string(c->block) + "/" +
                      std::to_string(force_block(c)) + "/" + std::to_string(c->D.P) + "/" +
                      std::to_string((int)any_erosion(c, -1));
    for (const Scene& s : c->scenes)
        key += "|" + std::to_string(s.ids.size()) + ":" + std::to_string(s.prm.h) + ":" +
               std::to_string(s.prm.box_half) + ":" + std::to_string(s.tnx) + "x" + std::to_string(s.tnz);
    Tuned best{c->yz, 1 << c->D.sub_log2, c->D.scan_log2};
    // every trial, and the context afterwards, starts from the state as it is now
    r = sync_to_host(c);
    if (r) return r;
    bool cached = false;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            best = it->second;
            cached = true;
        }
    }
    if (!cached) {
        r = collect_profile(c);
        if (r) return r;
        const bool prof = c->profiling;
        const auto saved_prof = c->prof;
        const Tuned orig = best;
        double best_ms = 1e30;
        // after a failed trial the grid goes back to what it was (best effort)
        auto restore = [&](int rc) {
            c->prof = saved_prof;
            c->profiling = prof;
            if (rc != SPHE_OK) {
                c->host_valid = true;
                (void)sphe_set_yz_split(c, orig.yz);
                (void)sphe_set_subcells(c, orig.slabs);
                c->scan_log2 = orig.scan_log2;
                if (c->layout_valid) (void)sphe_set_scan_tile(c, orig.scan_log2);
                c->device_valid = false;
            }
            return rc;
        };
        // the cell grid: y/z split and x slabs (the canonical order changes
        // with them; every combination is a tested configuration)
        for (int yz : {2, 1, 3}) {
            for (int slabs : {4, 2, 1}) {
                c->host_valid = true;  // a new split rebuilds the layout from the staged state
                r = sphe_set_yz_split(c, yz);
                if (r == SPHE_OK) r = prepare(c);
                if (r == SPHE_ERANGE || r == SPHE_ENOMEM) {
                    // a split with too many cells, or whose grid does not fit
                    // in memory, is not a candidate (the grid is rebuilt below)
                    (void)hipGetLastError();
                    break;
                }
                if (r == SPHE_OK) r = sphe_set_subcells(c, slabs);
                double ms = 0.0;
                if (r == SPHE_OK) r = grid_trial(c, steps_per_trial, &ms);
                if (r) return restore(r);
                if (ms < best_ms) {
                    best_ms = ms;
                    best.yz = yz;
                    best.slabs = slabs;
                }
            }
        }
        c->host_valid = true;
        r = sphe_set_yz_split(c, best.yz);
        if (r == SPHE_OK) r = prepare(c);
        if (r == SPHE_OK) r = sphe_set_subcells(c, best.slabs);
        if (r) return restore(r);
        // the scan tile on the chosen grid
        best_ms = 1e30;
        for (int t = SPHE_SCAN_LOG2_MIN; t <= SPHE_SCAN_LOG2_MAX; ++t) {
            r = sphe_set_scan_tile(c, t);
            double ms = 0.0;
            if (r == SPHE_OK) r = grid_trial(c, steps_per_trial, &ms);
            if (r) return restore(r);
            if (ms < best_ms) {
                best_ms = ms;
                best.scan_log2 = t;
            }
        }
        restore(SPHE_OK);
        std::lock_guard<std::mutex> g(mu);
        cache[key] = best;
    }
    c->host_valid = true;
    r = sphe_set_yz_split(c, best.yz);
    if (r == SPHE_OK) r = prepare(c);
    if (r == SPHE_OK) r = sphe_set_subcells(c, best.slabs);
    if (r == SPHE_OK) r = sphe_set_scan_tile(c, best.scan_log2);
    if (r) return r;
    c->host_valid = true;     // host staging still holds the pre-tuning state
    c->device_valid = false;  // restore it on the next use
    c->stepped = false;
    return SPHE_OK;
}

int sphe_debug_cell_keys(sphe_ctx* c, int scene, int32_t* keys)
{
    SCENE_OR_FAIL(s, c, scene);
    if (!keys) return fail(c, SPHE_EARG, "null output");
    int r = prepare(c);
    if (r) return r;
    const size_t n = s->ids.size();
    if (n == 0) return SPHE_OK;
    int32_t* dk = nullptr;
    HIPCHK(c, try_malloc(&dk, (size_t)s->p_cap * 4));
    sphe_launch_keys(c->D, scene, dk, c->stream);
    std::vector<int32_t> k(n), id(n);
    hipError_t e1 = hipMemcpyAsync(k.data(), dk, n * 4, hipMemcpyDeviceToHost, c->stream);
    hipError_t e2 = hipMemcpyAsync(id.data(), c->D.u_id + s->p_begin, n * 4, hipMemcpyDeviceToHost, c->stream);
    hipError_t e3 = hipStreamSynchronize(c->stream);
    (void)hipFree(dk);
    if (e1 || e2 || e3) return fail(c, SPHE_EHIP, "reading cell keys failed");
    std::vector<size_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return id[a] < id[b]; });
    for (size_t t = 0; t < n; ++t) keys[t] = k[perm[t]];
    return SPHE_OK;
}

int sphe_debug_consts(const sphe_ctx* c, int scene, double* out, int cap)
{
    const Scene* s = scene_of(c, scene);
    if (!s || !out || cap < 20) return SPHE_EARG;
    SceneDev d;
    std::memset(&d, 0, sizeof(d));
    derive(*s, d, c->yz);
    const double v[20] = {d.f.h, d.f.h2, d.f.thr, d.f.t3h2, d.f.mc6, d.f.pf_scale, d.f.vf_scale,
                          d.f.ng, d.f.h * 1.00390625f, d.f.inv_cell, d.f.org[0], d.f.org[1],
                          d.f.org[2], (double)d.dims[0], (double)d.dims[1], (double)d.dims[2],
                          d.hq, d.f.kcv, d.f.up, d.f.dn};
    std::memcpy(out, v, sizeof(v));
    return SPHE_OK;
}

int sphe_debug_neighbours(sphe_ctx* c, int scene, int32_t* counts, const int64_t* offsets, int32_t* idx)
{
    SCENE_OR_FAIL(s, c, scene);
    if (!counts) return fail(c, SPHE_EARG, "null counts");
    int r = sort_current(c);
    if (r) return r;
    const size_t n = s->ids.size();
    if (n == 0) return SPHE_OK;
    std::vector<int32_t> sid(n), cnt(n);
    HIPCHK(c, hipMemcpy(sid.data(), c->D.s_id + s->p_begin, n * 4, hipMemcpyDeviceToHost));
    int32_t* dcnt = nullptr;
    HIPCHK(c, try_malloc(&dcnt, n * 4));
    sphe_launch_neighbours(c->D, scene, nullptr, dcnt, nullptr, c->stream);
    hipError_t e = hipMemcpyAsync(cnt.data(), dcnt, n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        (void)hipFree(dcnt);
        return fail(c, SPHE_EHIP, "neighbour counts: %s", hipGetErrorString(e));
    }
    std::vector<size_t> perm(n);  // perm[t] = sorted slot of the t-th id
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return sid[a] < sid[b]; });
    for (size_t t = 0; t < n; ++t) counts[t] = cnt[perm[t]];
    if (!idx || !offsets) {
        (void)hipFree(dcnt);
        return SPHE_OK;
    }
    std::vector<int64_t> soff(n + 1, 0);
    for (size_t i = 0; i < n; ++i) soff[i + 1] = soff[i] + cnt[i];
    int64_t* doff = nullptr;
    int32_t* didx = nullptr;
    e = try_malloc(&doff, (n + 1) * 8);
    if (e == hipSuccess) e = try_malloc(&didx, (size_t)std::max<int64_t>(1, soff[n]) * 4);
    if (e == hipSuccess) e = hipMemcpy(doff, soff.data(), (n + 1) * 8, hipMemcpyHostToDevice);
    std::vector<int32_t> hidx((size_t)soff[n]);
    if (e == hipSuccess) {
        sphe_launch_neighbours(c->D, scene, doff, dcnt, didx, c->stream);
        e = hipMemcpyAsync(hidx.data(), didx, (size_t)soff[n] * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    (void)hipFree(dcnt);
    (void)hipFree(doff);
    (void)hipFree(didx);
    if (e != hipSuccess) return fail(c, SPHE_EHIP, "neighbour lists: %s", hipGetErrorString(e));
    for (size_t t = 0; t < n; ++t) {
        const size_t sl = perm[t];
        int32_t* dst = idx + offsets[t];
        std::memcpy(dst, hidx.data() + soff[sl], (size_t)cnt[sl] * 4);
        std::sort(dst, dst + cnt[sl]);
    }
    return SPHE_OK;
}

int sphe_debug_density(sphe_ctx* c, int scene, double* rho)
{
    SCENE_OR_FAIL(s, c, scene);
    if (!rho) return fail(c, SPHE_EARG, "null output");
    int r = sort_current(c);
    if (r) return r;
    c->stepped = true;  // sorted buffers and rho now describe the current state
    std::vector<double> v;
    r = read_rho(c, *s, v);
    c->stepped = false;
    if (r) return r;
    std::memcpy(rho, v.data(), v.size() * 8);
    return SPHE_OK;
}

int sphe_debug_accel(sphe_ctx* c, int scene, double* acc)
{
    SCENE_OR_FAIL(s, c, scene);
    if (!acc) return fail(c, SPHE_EARG, "null output");
    int r = sort_current(c);
    if (r) return r;
    const size_t n = s->ids.size();
    if (n == 0) return SPHE_OK;
    double* d = nullptr;
    HIPCHK(c, try_malloc(&d, n * 3 * 8));
    sphe_launch_accel(c->D, scene, c->lanes, d, c->stream);
    hipError_t e = hipGetLastError();
    std::vector<double> v(n * 3);
    std::vector<int32_t> id(n);
    if (e == hipSuccess) e = hipMemcpyAsync(v.data(), d, n * 3 * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(id.data(), c->D.s_id + s->p_begin, n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(c, SPHE_EHIP, "debug_accel: %s", hipGetErrorString(e));
    std::vector<size_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) { return id[a] < id[b]; });
    for (size_t t = 0; t < n; ++t)
        for (int a = 0; a < 3; ++a) acc[3 * t + a] = v[3 * perm[t] + a];
    return SPHE_OK;
}

int sphe_debug_copy(sphe_ctx* c, size_t bytes, int iters, double* avg_ms)
{
    if (!c || bytes == 0 || bytes % 16 || iters < 1) return fail(c, SPHE_EARG, "bytes must be a positive multiple of 16");
    void* a = nullptr;
    void* b = nullptr;
    HIPCHK(c, try_malloc(&a, bytes));
    hipError_t e = try_malloc(&b, bytes);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return fail(c, SPHE_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    hipEvent_t e0 = get_event(c), e1 = get_event(c);
    e = hipMemsetAsync(a, 0, bytes, c->stream);
    sphe_launch_copy(a, b, (int64_t)(bytes / 16), c->stream);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    for (int i = 0; i < iters && e == hipSuccess; ++i) sphe_launch_copy(a, b, (int64_t)(bytes / 16), c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    c->event_pool.push_back(e0);
    c->event_pool.push_back(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(c, SPHE_EHIP, "copy: %s", hipGetErrorString(e));
    if (avg_ms) *avg_ms = ms / iters;
    return SPHE_OK;
}

}  // extern "C"
