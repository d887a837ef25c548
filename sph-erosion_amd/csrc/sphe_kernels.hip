// sphe_kernels.hip -- CDNA4 (gfx950) kernels of the SPH-Erosion step.
//
// One step (all scenes of a context at once), reference Erosion/fluid_system.h
// FluidSystemSPH::Run (:104-183) + advance (:306-353):
//   scan      counts -> cell starts (2 kernels; tiles no particle counted
//             into are not read)
//   scatter   each particle to its cell segment (arrival order); resets
//             its cell's count for the next step
//   reorder   rank inside the cell by (x slab, particle id) -> canonical
//             (cell, slab, id) order; gathers the state into sorted SoA
//             buffers and writes the cell's slab offsets
//   density   poly6 density + pressure terms over the stencil rows clipped
//             to the smoothing sphere (fluid_system.h:109-124; cells are
//             h (1 + 2^-8) along x and half that along y and z, 25 rows),
//             and the neighbour lists for the force pass
//   force     pressure, viscosity, colour normal/Laplacian (:129-178),
//             symplectic Euler + box response (:318-350, :355-407),
//             heightfield collision and erosion exchange (library-defined),
//             and the cell count of the new position for the next step
//   heights   terrain heights from the fixed-point volumes
// The neighbour predicate is the reference's FP32 one, glm::length(r) <= h,
// evaluated exactly as d2 = (dx*dx + dy*dy) + dz*dz <= T(h) with T(h) the
// largest float whose square root is <= h.  Build with -ffp-contract=off.

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <map>
#include <tuple>
#include <algorithm>
#include <mutex>

#include "sphe_device.h"
#include "sphe_launch.h"

namespace {

template <typename T> struct V4s;
template <> struct V4s<float> { typedef float4 t; };
template <> struct V4s<double> { typedef double4 t; };
template <typename T> using V4 = typename V4s<T>::t;
template <typename T> struct V2s;
template <> struct V2s<float> { typedef float2 t; };
template <> struct V2s<double> { typedef double2 t; };
template <typename T> using V2 = typename V2s<T>::t;

__device__ __forceinline__ float d_floor(float x) { return __builtin_floorf(x); }
__device__ __forceinline__ double d_floor(double x) { return __builtin_floor(x); }
__device__ __forceinline__ float d_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double d_sqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float d_fabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double d_fabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float d_rint(float x) { return __builtin_rintf(x); }
__device__ __forceinline__ double d_rint(double x) { return __builtin_rint(x); }
__device__ __forceinline__ bool d_finite(float x) { return __builtin_isfinite(x); }
__device__ __forceinline__ bool d_finite(double x) { return __builtin_isfinite(x); }

// Blocks are dispatched round-robin over the 8 XCDs (block b on XCD b % 8).
// Chunked remap: XCD x takes runs of `chunk` consecutive tiles (run c of
// XCD x is tiles [(8c + x) chunk, (8c + x + 1) chunk)), so neighbouring
// cells share that XCD's L2 while the runs of the 8 XCDs stay interleaved
// over the domain (balanced when the work per particle varies).  chunk <= 1
// is the plain order; blocks past the last complete set of 8 runs keep
// their order.  Bijective for any count.
__device__ __forceinline__ int xcd_block(int bid, int nblk, int chunk)
{
    if (chunk <= 1) return bid;
    const int span = 8 * chunk;
    if (bid >= nblk / span * span) return bid;
    const int x = bid & 7, i = bid >> 3;
    return ((i / chunk) * 8 + x) * chunk + i % chunk;
}

__device__ __forceinline__ int wave_scene(const int32_t* chunk_scene, int64_t slot)
{
    return __builtin_amdgcn_readfirstlane(chunk_scene[slot >> 6]);
}

// Element j of a per-slot array, addressed by a 32-bit byte offset from the
// (wave-uniform) base: the layout keeps every such array under 4 GiB.
template <typename V>
__device__ __forceinline__ V at(const V* __restrict__ base, uint32_t j)
{
    return *(const V*)((const char*)base + j * (uint32_t)sizeof(V));
}

// The same at a byte offset (j * sizeof(V), precomputed).
template <typename V>
__device__ __forceinline__ V at_byte(const V* __restrict__ base, uint32_t off)
{
    return *(const V*)((const char*)base + off);
}

__device__ __forceinline__ float d_fmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double d_fmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float d_fmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double d_fmin(double a, double b) { return __builtin_fmin(a, b); }

// A floored cell coordinate clamped to [0, nmax]; negative and NaN give 0
// (max(NaN, 0) = 0), branch-free.  nmax is exact in T: the layout keeps grid
// and sub-cell extents below 2^24.
template <typename T>
__device__ __forceinline__ int cidx(T f, int nmax)
{
    return (int)d_fmin(d_fmax(f, T(0)), T(nmax));
}

template <typename T>
__device__ __forceinline__ void cell_of(const SceneConsts<T>& K, const int32_t* dims, T x, T y, T z,
                                        int& kx, int& ky, int& kz)
{
    kx = cidx<T>(d_floor((x - K.org[0]) * K.inv_cell), dims[0] - 1);
    ky = cidx<T>(d_floor((y - K.org[1]) * K.inv_yz), dims[1] - 1);
    kz = cidx<T>(d_floor((z - K.org[2]) * K.inv_yz), dims[2] - 1);
}

// Value of lane l ^ M.  Within a quad (M = 1, 2) a DPP quad permute (ALU
// only); otherwise a wave shuffle.
template <int M>
__device__ __forceinline__ float lane_xor(float v)
{
    if constexpr (M == 1) return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
    else if constexpr (M == 2) return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
    else return __shfl_xor(v, M, SPHE_WAVE);
}
template <int M>
__device__ __forceinline__ double lane_xor(double v)
{
    if constexpr (M <= 2) {
        const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
        constexpr int ctrl = M == 1 ? 0xB1 : 0x4E;  // quad_perm [1,0,3,2] / [2,3,0,1]
        const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)b, ctrl, 0xF, 0xF, false);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(b >> 32), ctrl, 0xF, 0xF, false);
        return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
    } else {
        return __shfl_xor(v, M, SPHE_WAVE);
    }
}

// Neighbour-sum accumulators.  A particle's sum is split over the G lanes of
// its group (density: lane l takes positions l, l+G, ... of the candidate
// stream; forces: list entries l, l+G, ...); merge_xor<M>() combines the
// partials of lanes l and l^M, so a xor tree over M = 1, 2, .. G/2 leaves
// the whole sum in every lane.
template <typename T, bool COMP> struct Acc;
template <typename T> struct Acc<T, true> {
    // Kahan: the represented sum is s - c
    T s = T(0), c = T(0);
    __device__ __forceinline__ void add(T v)
    {
        T y = v - c;
        T t = s + y;
        c = (t - s) - y;
        s = t;
    }
    // TwoSum of the heads (exact error e), compensations added: the result
    // is the same in both lanes of the pair.
    template <int M>
    __device__ __forceinline__ void merge_xor()
    {
        const T os = lane_xor<M>(s);
        const T oc = lane_xor<M>(c);
        const bool low = (threadIdx.x & M) == 0;
        const T a = low ? s : os, b = low ? os : s;
        const T ca = low ? c : oc, cb = low ? oc : c;
        const T sum = a + b;
        const T bb = sum - a;
        const T e = (a - (sum - bb)) + (b - bb);
        s = sum;
        c = (ca + cb) - e;
    }
    __device__ __forceinline__ T value() const { return s - c; }
};
template <typename T> struct Acc<T, false> {
    T s = T(0);
    __device__ __forceinline__ void add(T v) { s = s + v; }
    template <int M>
    __device__ __forceinline__ void merge_xor()
    {
        const T os = lane_xor<M>(s);
        const bool low = (threadIdx.x & M) == 0;
        s = low ? s + os : os + s;
    }
    __device__ __forceinline__ T value() const { return s; }
};

// Summation modes of the neighbour sums (the precision argument of
// sphe_create picks one): plain; Kahan-compensated density and force sums;
// compensated density sums with the force sums accumulated in FP64.
#define SUM_PLAIN 0
#define SUM_KAHAN 1
#define SUM_F64ACC 2

// FP64 sum of FP32 terms (plain adds; the lanes of a pair merge with the
// same commutative add, so both hold identical bits), rounded to FP32 once.
struct AccF64 {
    double s = 0.0;
    __device__ __forceinline__ void add(float v) { s = s + (double)v; }
    template <int M>
    __device__ __forceinline__ void merge_xor()
    {
        const double os = lane_xor<M>(s);
        s = s + os;
    }
    __device__ __forceinline__ float value() const { return (float)s; }
};

template <int G, typename A>
__device__ __forceinline__ void group_merge(A& a)
{
    if constexpr (G > 1) a.template merge_xor<1>();
    if constexpr (G > 2) a.template merge_xor<2>();
    if constexpr (G > 4) a.template merge_xor<4>();
}

// x sub-cell of a position inside its cell: the cell is split into
// 2^sub_log2 slabs along x, and a cell's particles are sorted by (slab, id).
// gx * 2^s is exact, so the slab index lies in [0, 2^s) for every position
// (clamped and non-finite ones included).
template <typename T>
__device__ __forceinline__ int sub_cell(T gx, int kx, int nx, int sl)
{
    return cidx<T>(d_floor(gx * T(1 << sl)), (nx << sl) - 1) - (kx << sl);
}

// A row gap g (a lower bound of the distance to a row, in x-cell units)
// less the rounding margin, floored at 0.
#define SPHE_ROW_MARGIN 0.015625
template <typename T>
__device__ __forceinline__ T row_gap(T f)
{
    const T a = f - T(SPHE_ROW_MARGIN);
    return a > T(0) ? a : T(0);
}

// The stencil rows of a particle, clipped to what the smoothing sphere can
// reach.  Cells are h (1 + 2^-8) along x and 1/yz of that along y and z
// (yz = 1, 2 or 3), so the rows are the (2 yz + 1)^2 (y, z) columns around the
// particle, numbered r = W (dz + yz) + (dy + yz), W = 2 yz + 1 (dz outer, dy
// inner).  In x-cell units the radius is h / cell < 1; a row whose y and z
// gaps g satisfy gy^2 + gz^2 > 1 is skipped, and the others keep the x
// sub-cells that meet [gx - w, gx + w], w = sqrt(1 - gy^2 - gz^2) + margin.
// Every neighbour is inside the range: the bounds are monotone in the
// position and include a margin far above the rounding of the cell
// coordinates.  Cells with more than 1023 particles carry no slab offsets
// and are taken whole.
// Half-width of a clipped row, in x-cell units, before the margin: at least
// sqrt(x) - 2e-3 for x in [0, 1].  FP64: sqrt.  FP32: x / sqrt(x) with one
// Newton step y <- y (1.5 - (x/2) y^2) for 1/sqrt(x) from the integer
// estimate 0x5f375a86 - bits(x)/2 (relative error at most 1.75e-3); the row
// margin (2^-6) covers the difference, and the sequence is IEEE operations
// only, so a CPU replays the same rows.
template <typename T> struct RowWidth {
    static __device__ __forceinline__ T of(T x) { return d_sqrt(x); }
};
template <> struct RowWidth<float> {
    static __device__ __forceinline__ float of(float x)
    {
        float y = __int_as_float(0x5f375a86 - (__float_as_int(x) >> 1));
        const float t = (0.5f * x) * y;
        y = y * __builtin_fmaf(-t, y, 1.5f);
        return x * y;
    }
};

template <typename T>
struct RowProbe {
    T gx, scale;
    // gaps, in x-cell units, to the rows one, two and three cells below (l1,
    // l2, l3) and above (h1, h2, h3) along y and z (the third with yz = 3)
    T yl1, yl2, yh1, yh2, zl1, zl2, zh1, zh2, yl3, yh3, zl3, zh3;
    int kx, ky, kz, sl, smax, fmax, flo_min, fhi_max;
    uint32_t rbase;  // global index of the first cell of the particle's own row
};

// r / W for 0 <= r < W^2 and W = 3, 5 or 7 (multiply and shift)
__host__ __device__ constexpr int row_div(int r, int W)
{
    return W == 3 ? (r * 11) >> 5 : (W == 5 ? (r * 13) >> 6 : (r * 37) >> 8);
}
constexpr bool row_div_exact()
{
    for (int W = 3; W <= 7; W += 2)
        for (int r = 0; r < W * W; ++r)
            if (row_div(r, W) != r / W) return false;
    return true;
}
static_assert(row_div_exact(), "row_div");

template <typename T>
__device__ __forceinline__ RowProbe<T> row_probe(const DevState& D, const SceneDev& S, const SceneConsts<T>& K,
                                                 T xi, T yi, T zi)
{
    RowProbe<T> P;
    const int nx = S.dims[0], ny = S.dims[1], nz = S.dims[2];
    P.gx = (xi - K.org[0]) * K.inv_cell;
    const T gy = (yi - K.org[1]) * K.inv_yz;
    const T gz = (zi - K.org[2]) * K.inv_yz;
    P.kx = cidx<T>(d_floor(P.gx), nx - 1);
    P.ky = cidx<T>(d_floor(gy), ny - 1);
    P.kz = cidx<T>(d_floor(gz), nz - 1);
    P.rbase = (uint32_t)S.cell_off + (uint32_t)((P.kz * ny + P.ky) * nx);
    const T fy = gy - T(P.ky), fz = gz - T(P.kz);
    const T q = K.yz_scale;  // 1 / yz, exact
    P.yl1 = row_gap<T>(fy * q);
    P.yh1 = row_gap<T>((T(1) - fy) * q);
    P.yl2 = row_gap<T>((T(1) + fy) * q);
    P.yh2 = row_gap<T>((T(1) + (T(1) - fy)) * q);
    P.zl1 = row_gap<T>(fz * q);
    P.zh1 = row_gap<T>((T(1) - fz) * q);
    P.zl2 = row_gap<T>((T(1) + fz) * q);
    P.zh2 = row_gap<T>((T(1) + (T(1) - fz)) * q);
    P.yl3 = P.yh3 = P.zl3 = P.zh3 = T(0);
    if (D.yz > 2) {  // uniform
        P.yl3 = row_gap<T>((T(2) + fy) * q);
        P.yh3 = row_gap<T>((T(2) + (T(1) - fy)) * q);
        P.zl3 = row_gap<T>((T(2) + fz) * q);
        P.zh3 = row_gap<T>((T(2) + (T(1) - fz)) * q);
    }
    P.sl = D.sub_log2;
    P.smax = (1 << P.sl) - 1;
    P.scale = T(1 << P.sl);
    P.fmax = (nx << P.sl) - 1;
    P.flo_min = (P.kx > 0 ? P.kx - 1 : 0) << P.sl;
    P.fhi_max = ((P.kx < nx - 1 ? P.kx + 1 : nx - 1) << P.sl) + P.smax;
    return P;
}

template <typename T>
__device__ __forceinline__ T gap_of(int d, int yz, T l1, T l2, T l3, T h1, T h2, T h3)
{
    // selects of values at hand: -1 l1, 1 h1, -2 l2, 2 h2, 0, and with
    // yz = 3 (a uniform branch) -3 l3, 3 h3
    const T one = d < 0 ? l1 : h1;
    const T two = d < 0 ? l2 : h2;
    T g = (d == 2 || d == -2) ? two : one;
    if (yz > 2) {
        const T three = d < 0 ? l3 : h3;
        g = (d == 3 || d == -3) ? three : g;
    }
    return d == 0 ? T(0) : g;
}

// Slot range [j0, j1) of the row at (dy, dz); false when the row is outside
// the grid or out of the sphere's reach.  roff = (dz ny + dy) nx, the row's
// cell offset from the particle's own row (mod 2^32), kept by the callers.  The six table words a row may need
// are loaded unconditionally (a skipped row reads the particle's own row),
// so a row costs one memory round trip.
template <typename T>
__device__ __forceinline__ bool row_range(const DevState& D, const SceneDev& S, const RowProbe<T>& P, int dyc,
                                          int dzc, uint32_t roff, uint32_t& j0, uint32_t& j1)
{
    const int ny = S.dims[1], nz = S.dims[2];
    const int z = P.kz + dzc, y = P.ky + dyc;
    const T mz = gap_of<T>(dzc, D.yz, P.zl1, P.zl2, P.zl3, P.zh1, P.zh2, P.zh3);
    const T my = gap_of<T>(dyc, D.yz, P.yl1, P.yl2, P.yl3, P.yh1, P.yh2, P.yh3);
    const T r2 = my * my + mz * mz;
    const bool on = (z >= 0) & (z < nz) & (y >= 0) & (y < ny) & (r2 <= T(1));
    const T w = RowWidth<T>::of(T(1) - (on ? r2 : T(0))) + T(SPHE_ROW_MARGIN);
    int flo = cidx<T>(d_floor((P.gx - w) * P.scale), P.fmax);
    int fhi = cidx<T>(d_floor((P.gx + w) * P.scale), P.fmax);
    flo = max(flo, P.flo_min);
    fhi = min(fhi, P.fhi_max);
    // cell indices are below 2^30 (the layout's limit): 32-bit offsets
    const uint32_t rk = on ? P.rbase + roff : P.rbase;
    const int sl = P.sl, smax = P.smax;
    const uint32_t clo = rk + (uint32_t)(flo >> sl), chi = rk + (uint32_t)(fhi >> sl);
    const uint32_t s_lo = at(D.start, clo), s_lo1 = at(D.start, clo + 1);
    const uint32_t s_hi = at(D.start, chi), s_hi1 = at(D.start, chi + 1);
    const uint32_t w0 = at(D.subw, clo), w1 = at(D.subw, chi);
    // slab offsets (none with one slab: slo = 0, shi = smax), clamped to the
    // cell: an empty cell's word is not rewritten each step, and any value
    // then gives its empty range; starts never decrease, so min(s_lo, s_lo1)
    // is s_lo when there is no offset
    const int slo = flo & smax, shi = fhi & smax;
    // field shifts from byte tables: 10 (slo - 1) for slo = 1..3, 10 shi
    const uint32_t sh0 = (0x140A0000u >> ((uint32_t)slo << 3)) & 0xFFu;
    const uint32_t sh1 = (0x1E140A00u >> ((uint32_t)shi << 3)) & 0xFFu;
    const uint32_t f0 = (slo != 0 && !(w0 >> 31)) ? (w0 >> sh0) & 1023u : 0u;
    const bool use1 = shi != smax && !(w1 >> 31);
    const uint32_t a = min(s_lo + f0, s_lo1);
    const uint32_t b = use1 ? min(s_hi + ((w1 >> sh1) & 1023u), s_hi1) : s_hi1;
    // rows of real cells lie inside [p_begin, p_begin + n); the clamp only
    // matters if the counts were inconsistent
    const uint32_t lo = (uint32_t)S.p_begin, hi = (uint32_t)(S.p_begin + S.n);
    j0 = max(a, lo);
    j1 = min(b, hi);
    return on;
}

// Calls row(j0, j1) for each row that may hold a neighbour, in row order.
template <typename T, typename F>
__device__ __forceinline__ void for_each_row(const DevState& D, const SceneDev& S, const SceneConsts<T>& K,
                                             T xi, T yi, T zi, F&& row)
{
    const RowProbe<T> P = row_probe<T>(D, S, K, xi, yi, zi);
    const int yz = D.yz;
    const int nx = S.dims[0], nyx = S.dims[1] * nx;
    for (int dz = -yz; dz <= yz; ++dz) {
        uint32_t roff = (uint32_t)(dz * nyx - yz * nx);
        for (int dy = -yz; dy <= yz; ++dy, roff += (uint32_t)nx) {
            uint32_t j0, j1;
            if (row_range<T>(D, S, P, dy, dz, roff, j0, j1)) row(j0, j1);
        }
    }
}

// Visits the candidates of the clipped stencil rows in ascending slot order
// and calls f(j, dx, dy, dz, d2) for those inside the smoothing radius.
template <typename T, typename F>
__device__ __forceinline__ void visit_neighbours(const DevState& D, const SceneDev& S, const SceneConsts<T>& K,
                                                 const V4<T>* __restrict__ pos, T xi, T yi, T zi, F&& f)
{
    for_each_row<T>(D, S, K, xi, yi, zi, [&](uint32_t j0, uint32_t j1) {
            uint32_t j = j0;
            for (; j + 1 < j1; j += 2) {  // two loads in flight
                const V4<T> pa = pos[j];
                const V4<T> pb = pos[j + 1];
                const T ax = xi - pa.x, ay = yi - pa.y, az = zi - pa.z;
                const T a2 = (ax * ax + ay * ay) + az * az;
                if (a2 <= K.thr) f(j, ax, ay, az, a2);
                const T bx = xi - pb.x, by = yi - pb.y, bz = zi - pb.z;
                const T b2 = (bx * bx + by * by) + bz * bz;
                if (b2 <= K.thr) f(j + 1, bx, by, bz, b2);
            }
            if (j < j1) {
                const V4<T> pj = pos[j];
                const T dx = xi - pj.x, dy = yi - pj.y, dz = zi - pj.z;
                const T d2 = (dx * dx + dy * dy) + dz * dz;
                if (d2 <= K.thr) f(j, dx, dy, dz, d2);
            }
    });
}

// Counting-sort key of a position: the cell key shifted left by sub_log2,
// or-ed with the x slab; the cell's count is taken with an atomic.
template <typename T>
__device__ __forceinline__ void count_position(const DevState& D, const SceneDev& S, const SceneConsts<T>& K,
                                               int64_t slot, T x, T y, T z)
{
    const T gx = (x - K.org[0]) * K.inv_cell;
    const int kx = cidx<T>(d_floor(gx), S.dims[0] - 1);
    const int ky = cidx<T>(d_floor((y - K.org[1]) * K.inv_yz), S.dims[1] - 1);
    const int kz = cidx<T>(d_floor((z - K.org[2]) * K.inv_yz), S.dims[2] - 1);
    const int64_t key = S.cell_off + ((int64_t)kz * S.dims[1] + ky) * S.dims[0] + kx;
    const int sub = sub_cell<T>(gx, kx, S.dims[0], D.sub_log2);
    D.key[slot] = (int32_t)((key << D.sub_log2) | sub);
    D.arank[slot] = atomicAdd(&D.cnt[key], 1u);
    D.tocc[key >> D.scan_log2] = 1u;
}

__device__ __forceinline__ void count_pad(const DevState& D, const SceneDev& S, int64_t slot)
{
    const int64_t key = S.cell_off + S.ncell;
    D.key[slot] = (int32_t)(key << D.sub_log2);
    D.arank[slot] = atomicAdd(&D.cnt[key], 1u);
    D.tocc[key >> D.scan_log2] = 1u;
}

// ------------------------------------------------------------------ sort

template <typename T>
__global__ __launch_bounds__(256) void k_hash_count(DevState D)
{
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= D.P) return;
    const SceneDev& S = D.scenes[wave_scene(D.chunk_scene, slot)];
    if (slot < S.p_begin + S.n) {
        const SceneConsts<T>& K = SceneSel<T>::get(S);
        const V4<T> p = ((const V4<T>*)D.u_pos)[slot];
        count_position<T>(D, S, K, slot, p.x, p.y, p.z);
    } else {
        count_pad(D, S, slot);
    }
}

// A scan tile is 256 threads x ITEMS consecutive counts (ITEMS = 8, 16 or
// 32: tiles of 2^11, 2^12 or 2^13 cells, DevState::scan_log2).

// Exclusive scan of v over a 256-thread block, with its total, and the
// block total of a second value w: one LDS exchange (lds: 8 words).
__device__ __forceinline__ uint32_t block_scan_sum(uint32_t v, uint32_t w, uint32_t* total, uint32_t* wtotal,
                                                   uint32_t* lds)
{
    // wave-inclusive scan and wave sum with shuffles, then the 4 wave totals
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    uint32_t r = w;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) r += __shfl_xor(r, o, 64);
    if (lane == 63) lds[wave] = x;
    if (lane == 0) lds[4 + wave] = r;
    __syncthreads();
    uint32_t base = 0, tot = 0, wt = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t t = lds[k];
        if (k < wave) base += t;
        tot += t;
        wt += lds[4 + k];
    }
    *total = tot;
    *wtotal = wt;
    return base + x - v;
}

// ITEMS consecutive words per thread, as 16-byte accesses where the whole
// run is inside [0, n) (tiles start at multiples of 64 bytes).
template <int ITEMS>
__device__ __forceinline__ void load_run(const uint32_t* __restrict__ p, int64_t base, int64_t n,
                                         uint32_t* v)
{
    if (base + ITEMS <= n) {
        const uint4* q = (const uint4*)(p + base);
#pragma unroll
        for (int k = 0; k < ITEMS / 4; ++k) {
            const uint4 a = q[k];
            v[4 * k] = a.x; v[4 * k + 1] = a.y; v[4 * k + 2] = a.z; v[4 * k + 3] = a.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) v[i] = base + i < n ? p[base + i] : 0u;
    }
}

template <int ITEMS>
__device__ __forceinline__ void store_run(uint32_t* __restrict__ p, int64_t base, int64_t n, const uint32_t* v)
{
    if (base + ITEMS <= n) {
        uint4* q = (uint4*)(p + base);
#pragma unroll
        for (int k = 0; k < ITEMS / 4; ++k) q[k] = make_uint4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
    } else {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
            if (base + i < n) p[base + i] = v[i];
    }
}

// Total of count tile `tile` into partial[tile].  occ (optional): the tile
// flags; an unflagged tile's counts are all zero.
template <int ITEMS>
__device__ __forceinline__ void scan_reduce_tile(const uint32_t* __restrict__ in, int64_t n,
                                                 uint32_t* __restrict__ partial, const uint32_t* __restrict__ occ,
                                                 int tile, uint32_t* lds)
{
    if (occ && occ[tile] == 0u) {
        if (threadIdx.x == 0) partial[tile] = 0u;
        return;
    }
    const int64_t base = (int64_t)tile * (256 * ITEMS) + (int64_t)threadIdx.x * ITEMS;
    uint32_t v[ITEMS];
    load_run<ITEMS>(in, base, n, v);
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) s += v[i];
    uint32_t tot, none;
    (void)block_scan_sum(0u, s, &none, &tot, lds);
    if (threadIdx.x == 0) partial[tile] = tot;
}

template <int ITEMS>
__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ in, int64_t n,
                                                     uint32_t* __restrict__ partial,
                                                     const uint32_t* __restrict__ occ)
{
    __shared__ uint32_t lds[8];
    scan_reduce_tile<ITEMS>(in, n, partial, occ, (int)blockIdx.x, lds);
}

// Exclusive prefix of the n tile totals, in place, by one block of 1,024
// threads: chunks of 16,384 totals, 16 consecutive ones per thread, one block
// scan per chunk and the running total carried to the next chunk.  Linear in
// the number of tiles; each scan-down block then reads its tile's prefix
// with one load.
__global__ __launch_bounds__(1024) void k_scan_tiles(uint32_t* __restrict__ partial, int64_t n)
{
    constexpr int ITEMS = 16;
    __shared__ uint32_t lds[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (int64_t c0 = 0; c0 < n; c0 += 1024 * ITEMS) {
        const int64_t base = c0 + (int64_t)threadIdx.x * ITEMS;
        uint32_t v[ITEMS];
        load_run<ITEMS>(partial, base, n, v);
        uint32_t s = 0;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) s += v[i];
        uint32_t x = s;  // inclusive scan over the wave, then the 16 wave totals
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) lds[wave] = x;
        __syncthreads();
        uint32_t before = 0, tot = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t t = lds[k];
            if (k < wave) before += t;
            tot += t;
        }
        uint32_t run = carry + before + (x - s);
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint32_t c = v[i];
            v[i] = run;
            run += c;
        }
        store_run<ITEMS>(partial, base, n, v);
        carry += tot;
        __syncthreads();  // the wave totals of the next chunk reuse lds
    }
}

// start[i] = exclusive prefix of counts, from partial[b]: the total of the
// tiles before tile b (k_scan_tiles).  With zero_counts the counts are
// reset to zero for the next arrivals (the particle sort instead has each
// particle's scatter reset its own cell); the optional single word *zero is
// set to zero_to.  occ (optional): the tile flags; an unflagged tile's counts are
// not read (all zero), and each block clears its tile's flag once read.
// fill (optional, with occ): per tile, the value all of an empty tile's
// starts hold since the tile was last written empty (~0: not known); an
// empty tile whose prefix equals it is not written again.
template <int ITEMS>
__global__ __launch_bounds__(256) void k_scan_down(uint32_t* __restrict__ cnt, uint32_t* __restrict__ start,
                                                   int64_t n, const uint32_t* __restrict__ partial,
                                                   int zero_counts, unsigned long long* __restrict__ zero,
                                                   unsigned long long zero_to,
                                                   uint32_t* __restrict__ occ, uint32_t* __restrict__ fill)
{
    if (zero && blockIdx.x == 0 && threadIdx.x == 0) *zero = zero_to;
    __shared__ uint32_t lds[8];
    const int64_t base = (int64_t)blockIdx.x * (256 * ITEMS) + (int64_t)threadIdx.x * ITEMS;
    const bool empty = occ && occ[blockIdx.x] == 0u;  // block-uniform
    const uint32_t filled = fill ? fill[blockIdx.x] : ~0u;
    uint32_t v[ITEMS];
    if (empty) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) v[i] = 0u;
    } else {
        load_run<ITEMS>(cnt, base, n, v);
    }
    // the total of the tiles before this one (k_scan_tiles' prefix)
    const uint32_t ptot = partial[blockIdx.x];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) s += v[i];
    uint32_t tot, none;
    uint32_t run = block_scan_sum(s, 0u, &tot, &none, lds);
    run += ptot;
    uint32_t z[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t c = v[i];
        v[i] = run;
        run += c;
        z[i] = 0u;
    }
    // an empty tile's starts all equal ptot (its prefix): unwritten when they
    // already hold it
    if (!(empty && filled == ptot)) store_run<ITEMS>(start, base, n, v);
    if (zero_counts && !empty) store_run<ITEMS>(cnt, base, n, z);
    // every thread read the flag and the fill word before the barrier of the
    // scan above
    if (threadIdx.x == 0) {
        if (occ && !empty) occ[blockIdx.x] = 0u;
        if (fill) {
            const uint32_t nf = empty ? ptot : ~0u;
            if (nf != filled) fill[blockIdx.x] = nf;
        }
    }
}

__global__ __launch_bounds__(256) void k_scatter(DevState D)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.P) return;
    const int32_t key = D.key[p];  // cell key << sub_log2 | slab
    const int32_t c = key >> D.sub_log2;
    if (key < 0 || (int64_t)c >= D.C) {
        atomicOr(D.status, 2u);
        return;
    }
    const uint32_t r = D.arank[p];
    const uint32_t d = D.start[c] + r;
    // the cell's count is used up: its first arrival (one per counted cell)
    // resets it for the next step's arrivals
    if (r == 0u) D.cnt[c] = 0u;
    if ((int64_t)d >= D.P) {  // counts inconsistent with the slots: report, never write out of range
        atomicOr(D.status, 2u);
        return;
    }
    // (key, id) as one word that compares as the pair: id's sign bit flipped
    D.tmp_kid[d] = ((uint64_t)(uint32_t)key << 32) | (uint32_t)(D.u_id[p] ^ (int32_t)0x80000000);
    D.tmp_src[d] = (int32_t)p;
}

// Canonical order inside a cell: ascending x slab, then ascending id (ties,
// which only duplicate ids produce, by previous slot).  The particle ranked
// first also writes the cell's slab offsets: for s = 1 .. 2^sub_log2 - 1 the
// number of the cell's particles in slabs below s, 10 bits each from bit
// 10 (s - 1); a cell of more than 1023 particles gets bit 31 instead.
// A particle's rank is counted over its cell's (key, id) words: the block
// first stages the words of the cells its 256 slots lie in (a window of at
// most SPHE_REORDER_WINDOW slots) in LDS, with one coalesced pass, so each
// word is read from memory about once; a block whose cells are larger ranks
// from memory.
#define SPHE_REORDER_WINDOW 2048
template <typename T>
__global__ __launch_bounds__(256) void k_reorder(DevState D)
{
    __shared__ uint64_t wkid[SPHE_REORDER_WINDOW];
    __shared__ uint32_t wlo, whi;
    const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = d < D.P;
    const int sl = D.sub_log2;
    uint64_t kid = 0;
    int32_t src = 0, c = 0;
    uint32_t s0 = 0, s1 = 0;
    bool ok = false;
    if (live) {
        kid = D.tmp_kid[d];
        src = D.tmp_src[d];
        const int32_t fk = (int32_t)(kid >> 32);
        c = fk >> sl;
        ok = fk >= 0 && (int64_t)c < D.C;
        if (ok) {
            s0 = D.start[c];
            s1 = (int64_t)D.start[c + 1] <= D.P ? D.start[c + 1] : (uint32_t)D.P;
        }
    }
    if (threadIdx.x == 0) {
        wlo = ~0u;
        whi = 0u;
    }
    __syncthreads();
    {
        // the block's window of cell slots: reduced over each wave, then one
        // LDS atomic pair per wave
        uint32_t lo = ok && s0 < s1 ? s0 : ~0u, hi = ok && s0 < s1 ? s1 : 0u;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const uint32_t a = __shfl_xor(lo, o, 64), b = __shfl_xor(hi, o, 64);
            lo = a < lo ? a : lo;
            hi = b > hi ? b : hi;
        }
        if ((threadIdx.x & 63) == 0 && lo < hi) {
            atomicMin(&wlo, lo);
            atomicMax(&whi, hi);
        }
    }
    __syncthreads();
    const uint32_t w0 = wlo, w1 = whi;
    const bool staged = w0 < w1 && w1 - w0 <= (uint32_t)SPHE_REORDER_WINDOW;  // block-uniform
    if (staged)
        for (uint32_t j = w0 + threadIdx.x; j < w1; j += blockDim.x) wkid[j - w0] = D.tmp_kid[j];
    __syncthreads();
    if (!live) return;
    if (!ok) {
        atomicOr(D.status, 2u);
        return;
    }
    // the particle's state, gathered while its rank is found (from a slot
    // clamped into range: a bad src is reported below and never stored)
    const uint32_t sc = (uint32_t)min(max(src, 0), (int32_t)(D.P - 1));
    const V4<T> pv = at((const V4<T>*)D.u_pos, sc);
    const V4<T> vv = at((const V4<T>*)D.u_vel, sc);
    const int64_t sed = D.u_sed[sc];
    const int32_t my = (int32_t)((uint32_t)kid ^ 0x80000000u);
    // one count field per slab boundary: a particle in slab k adds 1 to the
    // fields of the boundaries s > k
    const uint32_t fields = sl == 2 ? 0x100401u : (sl == 1 ? 1u : 0u);
    const uint32_t smask = (1u << sl) - 1u;
    uint32_t rank = 0, below = 0;
    auto take = [&](uint32_t j, uint64_t o) {
        rank += (o < kid || (o == kid && D.tmp_src[j] < src)) ? 1u : 0u;
        below += fields & (~0u << (10u * ((uint32_t)(o >> 32) & smask)));
    };
    uint32_t j = s0;
    if (staged) {
        for (; j + 3 < s1; j += 4) {
            const uint64_t o0 = wkid[j - w0], o1 = wkid[j + 1 - w0], o2 = wkid[j + 2 - w0], o3 = wkid[j + 3 - w0];
            take(j, o0);
            take(j + 1, o1);
            take(j + 2, o2);
            take(j + 3, o3);
        }
        for (; j < s1; ++j) take(j, wkid[j - w0]);
    } else {
        for (; j + 1 < s1; j += 2) {  // two loads in flight
            const uint64_t o0 = D.tmp_kid[j], o1 = D.tmp_kid[j + 1];
            take(j, o0);
            take(j + 1, o1);
        }
        if (j < s1) take(j, D.tmp_kid[j]);
    }
    const uint32_t f = s0 + rank;
    if ((int64_t)f >= D.P || src < 0 || (int64_t)src >= D.P) {
        atomicOr(D.status, 2u);
        return;
    }
    if (rank == 0) D.subw[c] = s1 - s0 > 1023u ? 0x80000000u : below;
    ((V4<T>*)D.s_pos)[f] = pv;
    ((V4<T>*)D.s_vel)[f] = vv;
    D.s_sed[f] = sed;
    D.s_id[f] = my;
}

// --------------------------------------------------------------- density

// Candidates per lane per iteration of the density pass (at most 4: one
// byte each in the placement word).
#define SPHE_DENSITY_UNROLL 4

// Inclusive prefix sum of x over the G lanes of a group (lane l gets
// x_0 + ... + x_l); *tot = the group's total.  The G lanes are consecutive
// lanes of the wave, starting at a multiple of G.
template <int G>
__device__ __forceinline__ uint32_t group_prefix(uint32_t x, int l, uint32_t* tot)
{
    uint32_t y = x;
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        const uint32_t o = __shfl_up(y, d, G);
        if (l >= d) y += o;
    }
    *tot = __shfl(y, G - 1, G);
    return y;
}

// Running sums of v's bytes 0..2 in bytes 0..2 of the result (b0, b0+b1,
// b0+b1+b2; no carries while they stay below 256): the 24-bit product with
// 0x010101, one full-rate v_mul_u32_u24.  Written as that instruction: the
// compiler otherwise folds a following shift and add into a 64-bit
// multiply-add, a quarter-rate instruction.
__device__ __forceinline__ uint32_t byte_runs(uint32_t v)
{
    uint32_t r;
    asm("v_mul_u32_u24_e32 %0, 0x10101, %1" : "=v"(r) : "v"(v));
    return r;
}

// G = 4: the groups are the wave's quads; three DPP quad permutes.
template <>
__device__ __forceinline__ uint32_t group_prefix<4>(uint32_t x, int l, uint32_t* tot)
{
    // quad_perm [0,0,1,2]: lane l reads lane l-1
    const uint32_t s1 = (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x90, 0xF, 0xF, false);
    uint32_t y = x + (l >= 1 ? s1 : 0u);
    // quad_perm [0,0,0,1]: lane l reads lane l-2
    const uint32_t s2 = (uint32_t)__builtin_amdgcn_mov_dpp((int)y, 0x40, 0xF, 0xF, false);
    y = y + (l >= 2 ? s2 : 0u);
    // quad_perm [3,3,3,3]: the total, from lane 3
    *tot = (uint32_t)__builtin_amdgcn_mov_dpp((int)y, 0xFF, 0xF, 0xF, false);
    return y;
}


// An LDS word addressed by its 32-bit LDS byte address (address space 3).
typedef __attribute__((address_space(3))) uint32_t lds_u32;

// Dynamic LDS of the density pass for ngroups lane groups: their row tables
// (W^2 + 1 entries of 8 bytes each), then their list staging rings (ring
// entries of 4 bytes each) from an offset aligned to a ring's size, so a
// ring slot's address is the ring's base or-ed with the slot's byte offset.
__host__ __device__ constexpr int density_ring_offset(int W, int ngroups, int ring)
{
    return (ngroups * (W * W + 1) * 8 + ring * 4 - 1) / (ring * 4) * (ring * 4);
}
__host__ __device__ constexpr int density_lds_bytes(int W, int ngroups, int ring)
{
    return density_ring_offset(W, ngroups, ring) + ngroups * ring * 4;
}

// One G-lane group per particle (lane l = thread index mod G).  The clipped
// stencil rows form one stream of candidates; lane l takes stream positions
// l, l+G, ... (U = SPHE_DENSITY_UNROLL per iteration) for its share of the
// density sum.
// The group writes one neighbour list per particle (self left out) in
// canonical order, nbl[slot * maxn + m], m < maxn, count in nbc[slot]; a
// byte-wise prefix sum over the group places each lane's entries in a ring
// in LDS, and every completed chunk of 16 entries goes out with one 16-byte
// store per lane (64-byte aligned: one cache line per group).  List
// entries are byte offsets of the neighbours' slots in the per-slot vectors
// of V4<T> (slot * sizeof(V4<T>), below 2^32 by the layout's limits).  The
// group then merges the partial sums and lane 0 writes rho and
// (p / rho^2, m / rho).
template <typename T, int SM, int G>
__device__ __forceinline__ void density_body(const DevState& D, int b)
{
    const int64_t gt = (int64_t)b * blockDim.x + threadIdx.x;
    const int64_t slot = gt / G;
    const int l = (int)(gt & (G - 1));
    if (slot >= D.P) return;
    const SceneDev& S = D.scenes[wave_scene(D.chunk_scene, slot)];
    if (slot >= S.p_begin + S.n) return;  // whole groups leave together
    const SceneConsts<T>& K = SceneSel<T>::get(S);
    const V4<T>* __restrict__ pos = (const V4<T>*)D.s_pos;
    const V4<T> pi = pos[slot];
    Acc<T, SM != SUM_PLAIN> sum;
    const T h2 = K.h2, thr = K.thr;
    const uint32_t cap = (uint32_t)D.maxn;
    uint32_t* __restrict__ nbl = D.nbl + slot * (int64_t)D.maxn;
    constexpr uint32_t VB = (uint32_t)sizeof(V4<T>);  // bytes per slot of a V4<T> vector
    const uint32_t self = (uint32_t)slot * VB;          // as a byte offset
    const int gfirst = (int)(threadIdx.x & (SPHE_WAVE - 1)) & ~(G - 1);  // the group's first lane in the wave
    uint32_t cnt = 0;      // the same in every lane of the group
    uint32_t pb = ~0u;     // the group's overflow chunk in the pool (none yet)
    bool claimed = false;  // the group asked for a chunk (at most once)
    uint32_t stream_total = 0;
    // The particle's clipped rows as one stream of candidates, t = 0 ..
    // total-1 in canonical order.  Lane l computes rows l, l+G, ... (the
    // group takes G rows per round) and writes (stream end, slot offset),
    // both in bytes (times VB), to the group's row table in LDS (dynamic:
    // blockDim / G tables of W^2 + 1 entries, then the groups' rings, laid
    // out by density_ring_offset): positions t of row r lie below
    // its end and hold slot t + offset.  A group prefix sum per round gives
    // the stream ends.  Entry W^2 is a sentinel row that ends at 2^32 - 1,
    // so a walk along the table stops by itself.
    const int yz = D.yz, W = 2 * yz + 1, nrows = W * W;
    extern __shared__ __attribute__((aligned(256))) uint2 rowtab[];
    constexpr int RING = SPHE_DENSITY_UNROLL * G > 16 ? 64 : 32;  // list staging ring (entries)
    constexpr uint32_t RMASK = RING * 4 - 1;                       // byte offset in a ring
    const int grp = (int)(threadIdx.x / G);
    uint2* __restrict__ tab = rowtab + grp * (nrows + 1);
    // the group's ring, size-aligned: slot offset ring_rel | (4 * position
    // mod RING) in the array; the hot stores use the LDS byte address
    // ring_abs | ..., the same with the array's own (256-byte aligned) address
    const uint32_t ring_rel = (uint32_t)(density_ring_offset(W, (int)(blockDim.x / G), RING) + grp * RING * 4);
    const uint32_t ring_abs = (uint32_t)(uintptr_t)rowtab + ring_rel;
    {
        const RowProbe<T> P = row_probe<T>(D, S, K, pi.x, pi.y, pi.z);
        uint32_t ends = 0;  // stream end of the rounds so far (all lanes)
        // row r = W dz + dy (dz, dy from 0): lane l starts at row l and steps
        // by G, carrying dz, dy and the row's cell offset (no multiplies)
        const int nx = S.dims[0], nyx = S.dims[1] * nx;
        int dz = row_div(l, W), dy = l - dz * W;
        uint32_t roff = (uint32_t)((dz - yz) * nyx + (dy - yz) * nx);
        const int rounds = (nrows + G - 1) / G;
        for (int k = 0, r = l; k < rounds; ++k, r += G) {
            uint32_t j0 = 0, len = 0;
            if (r < nrows) {
                uint32_t j1 = 0;
                const bool on = row_range<T>(D, S, P, dy - yz, dz - yz, roff, j0, j1);
                len = on && j1 > j0 ? j1 - j0 : 0u;
            }
            uint32_t rtot;
            const uint32_t end = ends + group_prefix<G>(len, l, &rtot);
            if (r < nrows) tab[r] = make_uint2(end * VB, (j0 - (end - len)) * VB);
            ends += rtot;
            dy += G;
            roff += (uint32_t)(G * nx);
            while (dy >= W) {
                dy -= W;
                ++dz;
                roff += (uint32_t)(nyx - W * nx);
            }
        }
        if (l == 0) tab[nrows] = make_uint2(~0u, 0u);
        // the group's lanes are one wave's: order its LDS writes before the reads
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        stream_total = ends;
    }
    // Lane l takes stream positions l, l+G, ... (U = SPHE_DENSITY_UNROLL
    // per iteration, blocks of U*G positions); its row pointer cur advances
    // through the table as its position passes each row's end (the sentinel
    // row ends every walk).  Positions, row ends and slots are in bytes
    // (times VB).  Blocks wholly inside the stream need no position checks;
    // in the last, partial block the positions past the stream load the
    // particle itself and take nothing.
    {
        constexpr int U = SPHE_DENSITY_UNROLL;
        const uint32_t total = stream_total;
        const uint2* cur = tab - 1;         // the current row's entry (none yet)
        uint32_t cur_end = 0, cur_off = 0;  // its stream end and slot offset
        const uint32_t lane_b = (uint32_t)l * VB;
        // Writes list positions c0 .. c0+n-1 (c0 a multiple of 16, n <= 16)
        // from the ring: lane l copies positions c0 + l PER .. + PER-1, as
        // 16-byte (G = 8: 8-byte) stores of whole chunks.  The list (capacity
        // cap, a multiple of 16) continues in the group's pool chunk of ocap
        // entries (a multiple of 16), claimed when the first chunk past cap is
        // written (one 64-bit atomic; none left: the rest is not stored, the
        // claim failure is counted and the force pass walks the stencil); the
        // list rows and pool chunks are 64-byte aligned.  Called by all lanes
        // of the group.
        const uint32_t ocap = (uint32_t)D.ocap;
        auto flush = [&](uint32_t c0, uint32_t n) {
            constexpr int PER = 16 / G;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            uint32_t* dst = nullptr;
            if (c0 < cap) {
                dst = nbl + c0;
            } else {
                if (!claimed) {
                    claimed = true;
                    uint32_t b0 = ~0u;
                    if (l == 0) {
                        const unsigned long long t = atomicAdd(D.ptop, (unsigned long long)ocap);
                        b0 = t + ocap <= (unsigned long long)D.pool_cap ? (uint32_t)(t / SPHE_POOL_UNIT) : ~0u;
                        if (b0 == ~0u) atomicAdd(&D.ovf[1], 1u);
                    }
                    pb = __shfl(b0, gfirst, SPHE_WAVE);
                }
                if (pb != ~0u && c0 - cap < ocap) dst = D.pool + (size_t)pb * SPHE_POOL_UNIT + (c0 - cap);
            }
            if (dst) {
                const uint32_t* src = (const uint32_t*)((char*)rowtab + (ring_rel | ((c0 * 4u) & RMASK))) + l * PER;
                uint32_t* out = dst + l * PER;
                if (n == 16u) {
                    if constexpr (PER >= 4) {
#pragma unroll
                        for (int k = 0; k < PER; k += 4) *(uint4*)(out + k) = *(const uint4*)(src + k);
                    } else {
                        *(uint2*)out = *(const uint2*)src;
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < PER; ++k)
                        if ((uint32_t)(l * PER + k) < n) out[k] = src[k];
                }
            }
            // the ring slots read here are written again only by later blocks
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        };
        auto block = [&](uint32_t base, auto tail) {
            constexpr bool TAIL = decltype(tail)::value;
            uint32_t jj[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t t = base + (uint32_t)(u * G + l);
                const uint32_t tb = base * VB + (uint32_t)(u * G) * VB + lane_b;
                while (tb >= cur_end) {
                    const uint2 e = *++cur;
                    cur_end = e.x;
                    cur_off = e.y;
                }
                ok[u] = TAIL ? t < total : true;
                jj[u] = ok[u] ? tb + cur_off : self;
            }
            T d2[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const V4<T> p = at_byte(pos, jj[u]);
                const T dx = pi.x - p.x, dy = pi.y - p.y, dz = pi.z - p.z;
                d2[u] = (dx * dx + dy * dy) + dz * dz;
            }
            // The lane's U terms (0 for a candidate outside the radius or
            // past the stream) are added in plain arithmetic, u = 0 first,
            // and the block's partial enters the lane's sum.  Outside the
            // radius the term is 0 without a test: d2 > thr >= h2 gives
            // q = 0 (h2 = fl(h h) <= thr since fl(sqrt(fl(h h))) = h).
            // byte u of x: 1 if this lane lists its u-th candidate
            uint32_t x = 0;
            T part = T(0);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool in = ok[u] & (d2[u] <= thr);
                T q = h2 - d2[u];
                q = q > T(0) ? q : T(0);
                const T w = ok[u] ? (q * q) * q : T(0);
                part = u ? part + w : w;
                // the force pass adds the particle's own term first: self is not listed
                x |= (uint32_t)(in & (jj[u] != self)) << (8 * u);
            }
            sum.add(part);
            // canonical place of the lane's u-th candidate (position u*G + l
            // of the block): the block's listed candidates at u' < u (all
            // lanes), then those of lanes below l at u.  Byte-wise sums over
            // the group (each byte stays <= 4 G, no carries): excl = lanes
            // below l, tot = all lanes; byte u of place is the sum of tot's
            // bytes below u plus excl's byte u.  sad_u8 adds tot's four bytes.
            uint32_t tot;
            const uint32_t excl = group_prefix<G>(x, l, &tot) - x;
            const uint32_t place = excl + (byte_runs(tot) << 8);
            const uint32_t ncnt = __builtin_amdgcn_sad_u8(tot, 0u, cnt);
            // the entries go to the group's ring in LDS (list position mod
            // RING, as byte offsets: place's bytes stay below 64, so times 4
            // they do not carry); each completed chunk of 16 is written out
            // at once
            const uint32_t place4 = place << 2, cnt4 = cnt << 2;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if ((x >> (8 * u)) & 1u)
                    *(lds_u32*)(uintptr_t)(ring_abs | ((cnt4 + ((place4 >> (8 * u)) & 0xFFu)) & RMASK)) = jj[u];
            // chunks completed by this block: from the one holding position
            // cnt (one at most when a block holds 16 positions or fewer)
            if constexpr (U * G <= 16) {
                const uint32_t c0 = cnt & ~15u;
                if (c0 + 16u <= ncnt) flush(c0, 16u);
            } else {
                for (uint32_t c0 = cnt & ~15u; c0 + 16u <= ncnt; c0 += 16u) flush(c0, 16u);
            }
            cnt = ncnt;
        };
        uint32_t base = 0;
        for (; base + U * G <= total; base += U * G) block(base, std::false_type());
        if (base < total) block(base, std::true_type());
        if (cnt & 15u) flush(cnt & ~15u, cnt & 15u);  // the last, partial chunk
    }
    group_merge<G>(sum);
    {
        // the largest count past the list: one atomic per wave (the lanes
        // still here; the loop over the overflowing groups is wave-uniform)
        uint64_t over = __ballot(cnt > cap);
        if (over) {
            uint32_t mx = 0;
            for (; over; over &= over - 1) {
                const uint32_t c = (uint32_t)__shfl(cnt, __ffsll((unsigned long long)over) - 1, SPHE_WAVE);
                mx = c > mx ? c : mx;
            }
            const uint64_t here = __ballot(1);
            if ((int)(threadIdx.x & (SPHE_WAVE - 1)) == __ffsll((unsigned long long)here) - 1) atomicMax(D.ovf, mx);
        }
    }
    if (l != 0) return;
    D.nbc[slot] = cnt;
    D.pbase[slot] = pb;
    const T rho = sum.value() * K.mc6;
    const T p = K.k * (rho - K.rho0);
    ((T*)D.rho)[slot] = rho;
    // (x, y, z, p / rho^2) and vel.w = m / rho: what the force pass gathers
    // per neighbour, in two 16-byte loads
    V4<T> pa = pi;
    pa.w = p / (rho * rho);
    ((V4<T>*)D.posa)[slot] = pa;
    ((T*)D.s_vel)[4 * slot + 3] = K.mass / rho;
}

template <typename T, int SM, int G>
__global__ __launch_bounds__(512) void k_density(DevState D, int nblk, int remap)
{
    density_body<T, SM, G>(D, xcd_block(blockIdx.x, nblk, remap));
}

// ----------------------------------------------------------------- force

template <typename T>
struct ForceOut {
    T pf[3], vf[3], nr[3], sf[3], gf[3], acc[3];
    int32_t nbmax;
    int32_t count;
};

// The terms one neighbour j != i adds to the ten force sums: pressure,
// viscosity, colour normal and colour Laplacian (fluid_system.h:136-150,
// 165-172) with the constants factored out.
// r = |d| and 1/r from d2 = |d|^2 (fluid_system.h:118, 438-441: length and
// normalize).  FP64: correctly rounded sqrt and division.  FP32: 1/sqrt by
// three Newton steps y <- y (1.5 - (d2/2) y^2) from the integer estimate
// 0x5f375a86 - bits(d2)/2 (relative error 3.4e-2 -> 1.8e-3 -> 4.6e-6 ->
// FP32 rounding), each step one multiply, one fused multiply-add and one
// multiply; r = d2 / sqrt(d2) = d2 y.  Within a few ulp of the correctly
// rounded values and made only of IEEE operations, so a CPU replays it bit
// for bit; it replaces two correctly rounded sequences of about ten
// instructions each.  d2 = 0 gives r = 0 (the coincident-pair branch).
template <typename T> struct Root {
    static __device__ __forceinline__ void of(T d2, T& r, T& inv)
    {
        r = d_sqrt(d2);
        inv = T(1) / r;
    }
};
template <> struct Root<float> {
    static __device__ __forceinline__ void of(float d2, float& r, float& inv)
    {
        float y = __int_as_float(0x5f375a86 - (__float_as_int(d2) >> 1));
        const float hx = 0.5f * d2;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            const float t = hx * y;
            y = y * __builtin_fmaf(-t, y, 1.5f);
        }
        inv = y;
        r = d2 * y;
    }
};

template <typename T> struct Terms {
    T lap, fp0, fp1, fp2, fv0, fv1, fv2, nn0, nn1, nn2;
};

// a + b, term by term (plain arithmetic)
template <typename T>
__device__ __forceinline__ Terms<T> terms_sum(const Terms<T>& a, const Terms<T>& b)
{
    Terms<T> t;
    t.lap = a.lap + b.lap;
    t.fp0 = a.fp0 + b.fp0; t.fp1 = a.fp1 + b.fp1; t.fp2 = a.fp2 + b.fp2;
    t.fv0 = a.fv0 + b.fv0; t.fv1 = a.fv1 + b.fv1; t.fv2 = a.fv2 + b.fv2;
    t.nn0 = a.nn0 + b.nn0; t.nn1 = a.nn1 + b.nn1; t.nn2 = a.nn2 + b.nn2;
    return t;
}

template <typename T>
__device__ __forceinline__ Terms<T> pair_terms(const SceneConsts<T>& K, T dx, T dy, T dz, T d2, T a_i, T a_j,
                                               T b_j, const V4<T>& vi, const V4<T>& vj)
{
    Terms<T> t;
    T q = K.h2 - d2;
    q = q > T(0) ? q : T(0);
    t.lap = b_j * (q * (K.t3h2 - T(7) * d2));
    T r, inv;
    Root<T>::of(d2, r, inv);
    const T hr = K.h - r;
    // coincident pair: the reference points along normalize(1,1,1)
    // (fluid_system.h:438-440); the reciprocal is discarded then
    const bool near = r < K.eps_r;
    const T ux = near ? K.inv_sqrt3 : dx * inv;
    const T uy = near ? K.inv_sqrt3 : dy * inv;
    const T uz = near ? K.inv_sqrt3 : dz * inv;
    const T sp = (a_i + a_j) * (hr * hr);
    t.fp0 = sp * ux;
    t.fp1 = sp * uy;
    t.fp2 = sp * uz;
    const T wv = b_j * hr;
    t.fv0 = (vj.x - vi.x) * wv;
    t.fv1 = (vj.y - vi.y) * wv;
    t.fv2 = (vj.z - vi.z) * wv;
    const T wn = b_j * (q * q);
    t.nn0 = dx * wn;
    t.nn1 = dy * wn;
    t.nn2 = dz * wn;
    return t;
}

// Force sums of one lane.  A lane's neighbours enter in consecutive groups
// of four, added in plain arithmetic as (a + b) + (c + d) before the group
// enters the lane's sum; the last one to three enter as a pair (a + b) and
// then a single.
// finish() applies the constants and forms the acceleration
// (fluid_system.h:151-154, 163, 177, 318-321).
// The force sums' accumulator: plain, Kahan-compensated, or (FP32 terms)
// FP64 sums (two instructions per term instead of the four of a
// compensated add) rounded to FP32 once.
template <typename T, int SM> struct ForceAccSel { typedef Acc<T, SM != SUM_PLAIN> type; };
template <> struct ForceAccSel<float, SUM_F64ACC> { typedef AccF64 type; };

template <typename T, int SM>
struct ForceAcc {
    typename ForceAccSel<T, SM>::type fp0, fp1, fp2, fv0, fv1, fv2, nn0, nn1, nn2, lap;

    // The particle's own colour-Laplacian term (d = 0): added first, by lane 0.
    __device__ __forceinline__ void self_term(const SceneConsts<T>& K, T b_i)
    {
        lap.add(b_i * (K.h2 * K.t3h2));
    }

    __device__ __forceinline__ void add(const Terms<T>& t)
    {
        lap.add(t.lap);
        fp0.add(t.fp0); fp1.add(t.fp1); fp2.add(t.fp2);
        fv0.add(t.fv0); fv1.add(t.fv1); fv2.add(t.fv2);
        nn0.add(t.nn0); nn1.add(t.nn1); nn2.add(t.nn2);
    }

    __device__ __forceinline__ void add2(const Terms<T>& a, const Terms<T>& b)
    {
        lap.add(a.lap + b.lap);
        fp0.add(a.fp0 + b.fp0); fp1.add(a.fp1 + b.fp1); fp2.add(a.fp2 + b.fp2);
        fv0.add(a.fv0 + b.fv0); fv1.add(a.fv1 + b.fv1); fv2.add(a.fv2 + b.fv2);
        nn0.add(a.nn0 + b.nn0); nn1.add(a.nn1 + b.nn1); nn2.add(a.nn2 + b.nn2);
    }

    // four consecutive terms as two plain pair sums ab = a + b, cd = c + d:
    // ab + cd enters the sums
    __device__ __forceinline__ void add4(const Terms<T>& ab, const Terms<T>& cd)
    {
        lap.add(ab.lap + cd.lap);
        fp0.add(ab.fp0 + cd.fp0); fp1.add(ab.fp1 + cd.fp1); fp2.add(ab.fp2 + cd.fp2);
        fv0.add(ab.fv0 + cd.fv0); fv1.add(ab.fv1 + cd.fv1); fv2.add(ab.fv2 + cd.fv2);
        nn0.add(ab.nn0 + cd.nn0); nn1.add(ab.nn1 + cd.nn1); nn2.add(ab.nn2 + cd.nn2);
    }

    template <int G>
    __device__ __forceinline__ void merge()
    {
        group_merge<G>(fp0); group_merge<G>(fp1); group_merge<G>(fp2);
        group_merge<G>(fv0); group_merge<G>(fv1); group_merge<G>(fv2);
        group_merge<G>(nn0); group_merge<G>(nn1); group_merge<G>(nn2);
        group_merge<G>(lap);
    }

    __device__ __forceinline__ void finish(const SceneConsts<T>& K, T rho_i, ForceOut<T>& o) const
    {
        const T sPF = K.pf_scale * rho_i;
        o.pf[0] = fp0.value() * sPF; o.pf[1] = fp1.value() * sPF; o.pf[2] = fp2.value() * sPF;
        o.vf[0] = fv0.value() * K.vf_scale; o.vf[1] = fv1.value() * K.vf_scale; o.vf[2] = fv2.value() * K.vf_scale;
        o.nr[0] = nn0.value() * K.ng; o.nr[1] = nn1.value() * K.ng; o.nr[2] = nn2.value() * K.ng;
        const T lapv = lap.value() * K.ng;
        const T ss = (-K.surf_tens) * lapv;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o.sf[a] = o.nr[a] * ss;
            o.gf[a] = rho_i * K.g[a];
            const T F = (o.pf[a] + o.vf[a]) + (o.gf[a] + o.sf[a]);
            o.acc[a] = F / rho_i;
        }
    }
};

// Lane l's share of the force sums of one particle by walking its clipped
// stencil rows (inspection, and particles whose list overflowed): every lane
// visits all candidates in canonical order and takes the accepted
// neighbours m = l, l+G, l+2G, ... (self excluded), as it would from the
// list.  count and nbmax cover the lane's share (self counted by lane 0).
template <typename T, int SM, int G>
__device__ __forceinline__ void force_share_walk(const DevState& D, const SceneDev& S, const SceneConsts<T>& K,
                                                 int64_t slot, int l, const V4<T>& pi, const V4<T>& vi, T a_i,
                                                 ForceAcc<T, SM>& acc, int32_t& nbmax, int32_t& count)
{
    const V4<T>* __restrict__ pos = (const V4<T>*)D.s_pos;
    const V4<T>* __restrict__ posa = (const V4<T>*)D.posa;
    const V4<T>* __restrict__ vel = (const V4<T>*)D.s_vel;
    // the start of a group of four a, b, c, d: ab = a (+ b), c
    Terms<T> ab, c;
    int np = 0;  // terms of the group so far
    uint32_t m = 0;    // canonical index of the next accepted neighbour
    if (l == 0) ++count;  // the particle itself
    visit_neighbours<T>(D, S, K, pos, pi.x, pi.y, pi.z, [&](uint32_t j, T dx, T dy, T dz, T d2) {
        if ((int64_t)j == slot) return;  // own term: ForceAcc::self_term
        const bool mine = (m & (G - 1)) == (uint32_t)l;
        ++m;
        if (!mine) return;
        ++count;
        const int32_t idj = D.s_id[j];
        nbmax = idj > nbmax ? idj : nbmax;
        const V4<T> vj = at(vel, j);
        const Terms<T> t = pair_terms<T>(K, dx, dy, dz, d2, a_i, at(posa, j).w, vj.w, vi, vj);
        if (np == 0) ab = t;
        else if (np == 1) ab = terms_sum(ab, t);
        else if (np == 2) c = t;
        else acc.add4(ab, terms_sum(c, t));
        np = (np + 1) & 3;
    });
    // the lane's last one to three terms: a, or the pair a + b (then c)
    if (np != 0) acc.add(ab);
    if (np == 3) acc.add(c);
}

// Dominant-axis box response (fluid_system.h:342-347 with collisionS
// :355-407), generalised to a cube of centre c and half extent e.
template <typename T>
__device__ __forceinline__ void box_response(const SceneConsts<T>& K, T* p, T* v)
{
    if ((K.box_lo[0] < p[0] && p[0] < K.box_hi[0]) && (K.box_lo[1] < p[1] && p[1] < K.box_hi[1]) &&
        (K.box_lo[2] < p[2] && p[2] < K.box_hi[2]))
        return;
    int ax = 0;
    T maxP = d_fabs(p[0] - K.box_c[0]);
    const T u1 = d_fabs(p[1] - K.box_c[1]);
    if (maxP < u1) { ax = 1; maxP = u1; }
    const T u2 = d_fabs(p[2] - K.box_c[2]);
    if (maxP < u2) { ax = 2; maxP = u2; }
    T cp[3] = {p[0], p[1], p[2]};
    T nr[3] = {T(0), T(0), T(0)};
    if (p[ax] < K.box_lo[ax]) { cp[ax] = K.box_lo[ax]; nr[ax] = T(1); }
    else { cp[ax] = K.box_hi[ax]; nr[ax] = T(-1); }
    const T d0 = p[0] - cp[0], d1 = p[1] - cp[1], d2 = p[2] - cp[2];
    const T d = d_sqrt((d0 * d0 + d1 * d1) + d2 * d2);
    const T vl = d_sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    const T s = T(1) + (T(0.5) * d) / (K.dt * vl);
    const T vn = (v[0] * nr[0] + v[1] * nr[1]) + v[2] * nr[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = v[a] - (nr[a] * s) * vn;
    p[0] = cp[0]; p[1] = cp[1]; p[2] = cp[2];
}

// Bilinear height of the cell-centred terrain at (x, z) and the unit normal
// of that patch (library-defined).
template <typename T>
__device__ __forceinline__ T terrain_sample(const SceneDev& S, const SceneConsts<T>& K, const float* __restrict__ H,
                                            T x, T z, T* nrm)
{
    const int tnx = S.tnx, tnz = S.tnz;
    const T u = (x - K.t_org[0]) * K.t_inv - T(0.5);
    const T w = (z - K.t_org[2]) * K.t_inv - T(0.5);
    const int i0 = cidx<T>(d_floor(u), tnx > 1 ? tnx - 2 : 0);
    const int k0 = cidx<T>(d_floor(w), tnz > 1 ? tnz - 2 : 0);
    const int i1 = tnx > 1 ? i0 + 1 : i0;
    const int k1 = tnz > 1 ? k0 + 1 : k0;
    T fu = u - T(i0), fw = w - T(k0);
    fu = fu > T(0) ? fu : T(0);
    fu = fu < T(1) ? fu : T(1);
    fw = fw > T(0) ? fw : T(0);
    fw = fw < T(1) ? fw : T(1);
    if (tnx == 1) fu = T(0);
    if (tnz == 1) fw = T(0);
    const float* Hs = H + S.t_off;
    const T h00 = (T)Hs[k0 * tnx + i0], h10 = (T)Hs[k0 * tnx + i1];
    const T h01 = (T)Hs[k1 * tnx + i0], h11 = (T)Hs[k1 * tnx + i1];
    const T gu = T(1) - fu, gw = T(1) - fw;
    const T a0 = gu * h00 + fu * h10;
    const T a1 = gu * h01 + fu * h11;
    const T hs = gw * a0 + fw * a1;
    const T dhu = gw * (h10 - h00) + fw * (h11 - h01);
    const T dhw = gu * (h01 - h00) + fu * (h11 - h10);
    const T gx = dhu * K.t_inv, gz = dhw * K.t_inv;
    const T nn = (gx * gx + T(1)) + gz * gz;
    const T inv = T(1) / d_sqrt(nn);
    nrm[0] = -gx * inv;
    nrm[1] = inv;
    nrm[2] = -gz * inv;
    return hs;
}

// Heightfield collision: projection onto the surface with the reference's
// response term (fluid_system.h:344-346) while moving into the terrain.
template <typename T>
__device__ __forceinline__ void terrain_collide(const SceneConsts<T>& K, T hs, const T* nr, T* p, T* v)
{
    if (!(p[1] < hs)) return;
    const T d = hs - p[1];
    const T vn = (v[0] * nr[0] + v[1] * nr[1]) + v[2] * nr[2];
    if (vn < T(0)) {
        const T vl = d_sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        const T s = T(1) + (T(0.5) * d) / (K.dt * vl);
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = v[a] - (nr[a] * s) * vn;
    }
    p[1] = hs;
}

// Erosion / deposition exchange of one particle (library-defined).  Returns
// the fixed-point amount taken from terrain cell *tc (negative: deposited),
// or 0 with *tc = -1 outside the contact band.
template <typename T>
__device__ __forceinline__ int64_t erode_exchange(const SceneDev& S, const SceneConsts<T>& K, const T* p, T hs,
                                                  const T* v, int64_t sed, int32_t* tc)
{
    *tc = -1;
    const T gap = p[1] - hs;
    if (!(gap <= K.band)) return 0;
    const int ti = cidx<T>(d_floor((p[0] - K.t_org[0]) * K.t_inv), S.tnx - 1);
    const int tk = cidx<T>(d_floor((p[2] - K.t_org[2]) * K.t_inv), S.tnz - 1);
    *tc = tk * S.tnx + ti;
    const T sp = d_sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    const T cap = K.kcv * sp;
    const T s = (T)sed * K.dn;
    T d;
    if (cap > s) d = K.ke * (cap - s);
    else d = -(K.kd * (s - cap));
    // one exchange saturates at 2^52 units (NaN: none), far inside int64
    if (!(d_fabs(d) <= K.dmax)) d = d > T(0) ? K.dmax : (d < T(0) ? -K.dmax : T(0));
    int64_t dq = (int64_t)d_rint(d * K.up);
    if (dq < -sed) dq = -sed;
    return dq;
}

// Lane l's share of the force sums from list entries l, l+G, l+2G, ... of
// a particle's neighbour sequence (fetch(k) = the byte offset of its k-th
// neighbour's slot in the V4<T> vectors): the lanes
// of a particle differ by one neighbour at most.  Four neighbours per
// iteration (eight gathers in flight), summed (a + b) + (c + d) before they
// enter the lane's sums; then one pair, then a single.
template <typename T, int SM, int G, typename F>
__device__ __forceinline__ void force_list(const DevState& D, const SceneConsts<T>& K, int l, uint32_t cnt,
                                          const V4<T>& pi, const V4<T>& vi, T a_i, ForceAcc<T, SM>& acc,
                                          F&& fetch)
{
    const V4<T>* __restrict__ posa = (const V4<T>*)D.posa;
    const V4<T>* __restrict__ vel = (const V4<T>*)D.s_vel;
    uint32_t k = (uint32_t)l;
    // the list entries of the next four-neighbour iteration load while the
    // current one computes
    uint32_t na = 0, nb = 0, nc = 0, nd = 0;
    if (k + 3 * G < cnt) { na = fetch(k); nb = fetch(k + G); nc = fetch(k + 2 * G); nd = fetch(k + 3 * G); }
    for (; k + 3 * G < cnt; k += 4 * G) {
        const uint32_t ja = na, jb = nb, jc = nc, jd = nd;
        if (k + 7 * G < cnt) {
            na = fetch(k + 4 * G); nb = fetch(k + 5 * G); nc = fetch(k + 6 * G); nd = fetch(k + 7 * G);
        }
        const V4<T> pa = at_byte(posa, ja), va = at_byte(vel, ja), pb = at_byte(posa, jb), vb = at_byte(vel, jb);
        const V4<T> pc = at_byte(posa, jc), vc = at_byte(vel, jc), pd = at_byte(posa, jd), vd = at_byte(vel, jd);
        Terms<T> ab, cd;
        {
            const T ax = pi.x - pa.x, ay = pi.y - pa.y, az = pi.z - pa.z;
            const T a2 = (ax * ax + ay * ay) + az * az;
            const T bx = pi.x - pb.x, by = pi.y - pb.y, bz = pi.z - pb.z;
            const T b2 = (bx * bx + by * by) + bz * bz;
            ab = terms_sum(pair_terms<T>(K, ax, ay, az, a2, a_i, pa.w, va.w, vi, va),
                           pair_terms<T>(K, bx, by, bz, b2, a_i, pb.w, vb.w, vi, vb));
        }
        {
            const T ax = pi.x - pc.x, ay = pi.y - pc.y, az = pi.z - pc.z;
            const T a2 = (ax * ax + ay * ay) + az * az;
            const T bx = pi.x - pd.x, by = pi.y - pd.y, bz = pi.z - pd.z;
            const T b2 = (bx * bx + by * by) + bz * bz;
            cd = terms_sum(pair_terms<T>(K, ax, ay, az, a2, a_i, pc.w, vc.w, vi, vc),
                           pair_terms<T>(K, bx, by, bz, b2, a_i, pd.w, vd.w, vi, vd));
        }
        acc.add4(ab, cd);
    }
    for (; k + G < cnt; k += 2 * G) {
        const uint32_t ja = fetch(k);
        const uint32_t jb = fetch(k + G);
        const V4<T> pa = at_byte(posa, ja);
        const V4<T> va = at_byte(vel, ja);
        const V4<T> pb = at_byte(posa, jb);
        const V4<T> vb = at_byte(vel, jb);
        const T ax = pi.x - pa.x, ay = pi.y - pa.y, az = pi.z - pa.z;
        const T a2 = (ax * ax + ay * ay) + az * az;
        const T bx = pi.x - pb.x, by = pi.y - pb.y, bz = pi.z - pb.z;
        const T b2 = (bx * bx + by * by) + bz * bz;
        acc.add2(pair_terms<T>(K, ax, ay, az, a2, a_i, pa.w, va.w, vi, va),
                 pair_terms<T>(K, bx, by, bz, b2, a_i, pb.w, vb.w, vi, vb));
    }
    if (k < cnt) {
        const uint32_t j = fetch(k);
        const V4<T> pj = at_byte(posa, j);
        const V4<T> vj = at_byte(vel, j);
        const T dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const T d2 = (dx * dx + dy * dy) + dz * dz;
        acc.add(pair_terms<T>(K, dx, dy, dz, d2, a_i, pj.w, vj.w, vi, vj));
    }
}

// Lane l's share of the force sums of one particle: from its list, the
// list continued in its overflow chunk, or (no room for all neighbours)
// by walking the stencil.
template <typename T, int SM, int G>
__device__ __forceinline__ void force_share(const DevState& D, const SceneDev& S, const SceneConsts<T>& K,
                                           int64_t slot, int l, const V4<T>& pi, const V4<T>& vi, T a_i,
                                           ForceAcc<T, SM>& acc, bool report_walk)
{
    const uint32_t cnt = D.nbc[slot];
    const uint32_t cap = (uint32_t)D.maxn;
    const uint32_t* __restrict__ nbl = D.nbl + slot * (int64_t)D.maxn;
    if (cnt <= cap) {
        force_list<T, SM, G>(D, K, l, cnt, pi, vi, a_i, acc, [&](uint32_t k) { return nbl[k]; });
        return;
    }
    const uint32_t pb = D.pbase[slot];
    if (pb != ~0u && cnt - cap <= (uint32_t)D.ocap) {
        const uint32_t* __restrict__ chunk = D.pool + (size_t)pb * SPHE_POOL_UNIT;
        force_list<T, SM, G>(D, K, l, cnt, pi, vi, a_i, acc,
                               [&](uint32_t k) { return k < cap ? nbl[k] : chunk[k - cap]; });
        return;
    }
    // no room in list and pool: the stencil again (same order, same result)
    if (report_walk && l == 0) atomicOr(D.status, 4u);
    int32_t nbmax = -1, count = 0;
    force_share_walk<T, SM, G>(D, S, K, slot, l, pi, vi, a_i, acc, nbmax, count);
}

// One G-lane group per particle: the lanes sum their shares, merge them
// (xor shuffles), and lane 0 integrates, collides, erodes, stores the new
// state and counts the particle's next cell.
template <typename T, int SM, int G>
__device__ __forceinline__ void force_body(const DevState& D, int b)
{
    const int64_t gt = (int64_t)b * blockDim.x + threadIdx.x;
    const int64_t slot = gt / G;
    const int l = (int)(gt & (G - 1));
    if (slot >= D.P) return;
    const SceneDev& S = D.scenes[wave_scene(D.chunk_scene, slot)];
    V4<T>* __restrict__ upos = (V4<T>*)D.u_pos;
    V4<T>* __restrict__ uvel = (V4<T>*)D.u_vel;
    const V4<T>* __restrict__ spos = (const V4<T>*)D.s_pos;
    const V4<T>* __restrict__ svel = (const V4<T>*)D.s_vel;
    if (slot >= S.p_begin + S.n) {
        if (l != 0) return;
        // padding: carried through, counted into the scene's pad cell
        upos[slot] = spos[slot];
        uvel[slot] = svel[slot];
        D.u_sed[slot] = D.s_sed[slot];
        D.u_id[slot] = D.s_id[slot];
        count_pad(D, S, slot);
        if (S.erosion == 2) D.tcell[slot] = -1;
        return;
    }
    const SceneConsts<T>& K = SceneSel<T>::get(S);
    const V4<T> pi = spos[slot];
    const V4<T> vi = svel[slot];
    const T rho_i = ((const T*)D.rho)[slot];
    const T a_i = ((const V4<T>*)D.posa)[slot].w;
    ForceAcc<T, SM> acc;
    if (l == 0) acc.self_term(K, vi.w);
    force_share<T, SM, G>(D, S, K, slot, l, pi, vi, a_i, acc, true);
    acc.template merge<G>();
    if (l != 0) return;
    ForceOut<T> o;
    acc.finish(K, rho_i, o);

    // symplectic Euler (fluid_system.h:330-331)
    T p[3], v[3];
    const T xin[3] = {pi.x, pi.y, pi.z}, vin[3] = {vi.x, vi.y, vi.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        v[a] = vin[a] + o.acc[a] * K.dt;
        p[a] = xin[a] + v[a] * K.dt;
    }
    int64_t sed = D.s_sed[slot];
    const bool moving = K.dt != T(0);
    if (moving) box_response<T>(K, p, v);
    int32_t tc = -1;
    int64_t dq = 0;
    if (S.tnx > 0) {
        T nr[3];
        const T hs = terrain_sample<T>(S, K, D.height, p[0], p[2], nr);
        if (moving) {
            terrain_collide<T>(K, hs, nr, p, v);
            if (S.erosion) dq = erode_exchange<T>(S, K, p, hs, v, sed, &tc);
        }
    }
    if (S.erosion == 1 && dq != 0) {
        sed += dq;
        atomicAdd((unsigned long long*)&D.vol[S.t_off + tc], (unsigned long long)(-dq));
    } else if (S.erosion == 2) {
        sed += dq;
        D.dq[slot] = dq;
        D.tcell[slot] = dq != 0 ? S.t_off + tc : -1;
    }
    V4<T> np, nv;
    np.x = p[0]; np.y = p[1]; np.z = p[2]; np.w = T(0);
    nv.x = v[0]; nv.y = v[1]; nv.z = v[2]; nv.w = T(0);
    upos[slot] = np;
    uvel[slot] = nv;
    D.u_sed[slot] = sed;
    D.u_id[slot] = D.s_id[slot];
    if (!(d_finite(p[0]) && d_finite(p[1]) && d_finite(p[2]) && d_finite(v[0]) && d_finite(v[1]) &&
          d_finite(v[2])))
        atomicOr(D.status, 1u);
    // cell count of the new position for the next step's sort
    count_position<T>(D, S, K, slot, p[0], p[1], p[2]);
}

// FP32: at most 96 registers (5 waves per SIMD): the list loop fits, and
// only the rarely taken stencil walk of overflowing particles spills.
// Unbounded, the walk raises the count and the occupancy of the whole
// kernel drops to 4.
template <int SM, int G>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_force_f32(DevState D, int nblk,
                                                                                              int remap)
{
    force_body<float, SM, G>(D, xcd_block(blockIdx.x, nblk, remap));
}

// FP64: at most 168 registers (3 waves per SIMD instead of 2 unbounded).
template <int G>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_force_f64(DevState D, int nblk,
                                                                                              int remap)
{
    force_body<double, SUM_PLAIN, G>(D, xcd_block(blockIdx.x, nblk, remap));
}

// ------------------------------------------------------ persistent passes

// Persistent form of the density and force passes: a grid of resident
// blocks takes the passes' blocks (work items, in the flattened slot order
// of all scenes) from eight queues, one per XCD: block p starts on queue
// p mod 8 (the XCD the hardware gives it), whose k-th item is the one the
// ordinary grid runs as block 8 k + x, remapped in runs of `chunk` as
// xcd_block does; a block whose queue is empty takes items from the other
// queues.  A block takes SPHE_GRAB consecutive items of a queue with one
// atomic and asks for the next ones while it works on them.  Every item is
// taken exactly once, and a block leaves when all eight queues are empty,
// so every wave exits.  The queues (work[0..7], in items) are cleared by
// k_work_reset before each launch.  The items of one block share no LDS
// between waves (each lane group has its own region), so one barrier per
// grab suffices.
#define SPHE_GRAB 4
__device__ __forceinline__ int64_t next_grab(uint32_t* work, int x, int nblk)
{
    for (int t = 0; t < 8; ++t) {
        const int q = (x + t) & 7;
        if ((int64_t)__hip_atomic_load(&work[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) * 8 + q >= nblk)
            continue;  // drained
        const int64_t bid = (int64_t)atomicAdd(&work[q], (uint32_t)SPHE_GRAB) * 8 + q;
        if (bid < nblk) return bid;
    }
    return -1;
}

__global__ void k_work_reset(uint32_t* work)
{
    if (threadIdx.x < 8) work[threadIdx.x] = 0u;
}

// g: the first item of a grab as an ordinary block index 8 k + q
#define SPHE_PERSISTENT_LOOP(BODY)                                                   \
    __shared__ int64_t s_grab[2];                                                    \
    const int x = (int)(blockIdx.x & 7u);                                            \
    if (threadIdx.x == 0) s_grab[0] = next_grab(work, x, nblk);                      \
    __syncthreads();                                                                 \
    int64_t g = s_grab[0];                                                           \
    for (int k = 0; g >= 0; ++k) {                                                   \
        const int64_t nxt = threadIdx.x == 0 ? next_grab(work, x, nblk) : -1;        \
        for (int i = 0; i < SPHE_GRAB; ++i) {                                        \
            const int64_t bid = g + 8 * i;                                           \
            if (bid >= nblk) break;                                                  \
            const int b = xcd_block((int)bid, nblk, chunk);                          \
            BODY;                                                                    \
        }                                                                            \
        if (threadIdx.x == 0) s_grab[(k + 1) & 1] = nxt;                             \
        __syncthreads();                                                             \
        g = s_grab[(k + 1) & 1];                                                     \
    }

template <typename T, int SM, int G>
__global__ __launch_bounds__(512) void k_density_pers(DevState D, int nblk, int chunk, uint32_t* work)
{
    SPHE_PERSISTENT_LOOP((density_body<T, SM, G>(D, b)))
}

template <int SM, int G>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_force_f32_pers(
    DevState D, int nblk, int chunk, uint32_t* work)
{
    SPHE_PERSISTENT_LOOP((force_body<float, SM, G>(D, b)))
}

template <int G>
__global__ __launch_bounds__(512) void k_force_f64_pers(DevState D, int nblk, int chunk, uint32_t* work)
{
    SPHE_PERSISTENT_LOOP((force_body<double, SUM_PLAIN, G>(D, b)))
}

// --------------------------------------------------------------- terrain

// Terrain heights from the fixed-point volumes, four consecutive cells per
// thread (16-byte loads and store).  A scene's range is padded to 64 cells,
// so the four cells share a scene; cells past the scene's nx * nz are
// padding and keep their value.
__device__ __forceinline__ void heights_quad(const DevState& D, int64_t t0)
{
    const SceneDev& S = D.scenes[D.tchunk_scene[t0 >> 6]];
    const int64_t end = S.t_off + (int64_t)S.tnx * S.tnz;
    const float y0 = S.t_y0, hq = S.hq;
    if (t0 + 4 <= end) {
        const longlong2 a = *(const longlong2*)(D.vol + t0);
        const longlong2 b = *(const longlong2*)(D.vol + t0 + 2);
        *(float4*)(D.height + t0) = make_float4(y0 + (float)a.x * hq, y0 + (float)a.y * hq,
                                                y0 + (float)b.x * hq, y0 + (float)b.y * hq);
    } else {
        for (int64_t t = t0; t < end; ++t) D.height[t] = y0 + (float)D.vol[t] * hq;
    }
}

__global__ __launch_bounds__(256) void k_heights(DevState D)
{
    const int64_t t0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (t0 < D.TC) heights_quad(D, t0);
}

// The particle scan's first kernel and the terrain heights in one launch
// (both need only the force pass done): blocks [0, nscan) total the count
// tiles, the others compute heights from the volumes the last force pass
// eroded, before this step's force pass samples them.
template <int ITEMS>
__global__ __launch_bounds__(256) void k_scan_reduce_heights(DevState D, int nscan)
{
    __shared__ uint32_t lds[8];
    if ((int)blockIdx.x >= nscan) {  // block-uniform: no barrier is skipped by part of a block
        const int64_t t0 = ((int64_t)(blockIdx.x - nscan) * blockDim.x + threadIdx.x) * 4;
        if (t0 < D.TC) heights_quad(D, t0);
        return;
    }
    scan_reduce_tile<ITEMS>(D.cnt, D.C + 1, D.partial, D.tocc, (int)blockIdx.x, lds);
}

// Two-phase erosion, phase 2a: count exchanges per terrain cell.
__global__ __launch_bounds__(256) void k_tcount(DevState D)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.P) return;
    const SceneDev& S = D.scenes[wave_scene(D.chunk_scene, p)];
    if (S.erosion != 2) return;
    const int32_t tc = D.tcell[p];
    if (tc >= 0) atomicAdd(&D.tcnt[tc], 1u);
}

// Phase 2b: place each exchange in its terrain cell's segment.
__global__ __launch_bounds__(256) void k_tscatter(DevState D)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.P) return;
    const SceneDev& S = D.scenes[wave_scene(D.chunk_scene, p)];
    if (S.erosion != 2) return;
    const int32_t tc = D.tcell[p];
    if (tc < 0) return;
    const uint32_t at = D.tstart[tc] + atomicAdd(&D.tcnt[tc], 1u);
    D.tdq[at] = D.dq[p];
}

// Phase 2c: one read-modify-write per terrain cell; the integer sum does not
// depend on the order the segment was filled in.
__global__ __launch_bounds__(256) void k_tapply(DevState D)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.TC) return;
    const uint32_t s0 = D.tstart[t], s1 = D.tstart[t + 1];
    int64_t sum = 0;
    for (uint32_t j = s0; j < s1; ++j) sum += D.tdq[j];
    if (s1 > s0) D.vol[t] -= sum;
    D.tcnt[t] = 0u;
}

// ----------------------------------------------------------- inspection

// The record of one particle after a step: its group recomputes the force
// sums of that step (same lane split and merge as k_force) from the retained
// sorted state, and reads the state the step left in the same slot.  The
// slot comes from device memory (the id map or the id search; INT32_MAX or
// out of range: no such particle, out[22] = -1).
template <typename T, int SM, int G>
__global__ void k_inspect(DevState D, const int32_t* __restrict__ slotp, double* out)
{
    if (blockIdx.x != 0 || threadIdx.x >= G) return;
    const int32_t s32 = *slotp;
    if (s32 < 0 || (int64_t)s32 >= D.P) {  // the same in every lane
        if (threadIdx.x == 0) out[22] = -1.0;
        return;
    }
    const int64_t slot = s32;
    const int l = (int)threadIdx.x;
    const SceneDev& S = D.scenes[D.chunk_scene[slot >> 6]];
    const SceneConsts<T>& K = SceneSel<T>::get(S);
    const V4<T> pi = ((const V4<T>*)D.s_pos)[slot];
    const V4<T> vi = ((const V4<T>*)D.s_vel)[slot];
    const T rho_i = ((const T*)D.rho)[slot];
    const T a_i = ((const V4<T>*)D.posa)[slot].w;
    ForceAcc<T, SM> acc;
    if (l == 0) acc.self_term(K, vi.w);
    int32_t nbmax = -1, count = 0;
    force_share_walk<T, SM, G>(D, S, K, slot, l, pi, vi, a_i, acc, nbmax, count);
    acc.template merge<G>();
#pragma unroll
    for (int m = 1; m < G; m <<= 1) {
        nbmax = max(nbmax, __shfl_xor(nbmax, m, SPHE_WAVE));
        count += __shfl_xor(count, m, SPHE_WAVE);
    }
    if (l != 0) return;
    ForceOut<T> o;
    acc.finish(K, rho_i, o);
    o.nbmax = nbmax;
    o.count = count;
    out[0] = (double)rho_i;
    out[1] = (double)(K.k * (rho_i - K.rho0));
    for (int a = 0; a < 3; ++a) {
        out[2 + a] = (double)o.acc[a];
        out[5 + a] = (double)o.pf[a];
        out[8 + a] = (double)o.vf[a];
        out[11 + a] = (double)o.gf[a];
        out[14 + a] = (double)o.sf[a];
        out[17 + a] = (double)o.nr[a];
    }
    out[20] = (double)o.nbmax;
    out[21] = (double)o.count;
    out[22] = (double)slot;
    const V4<T> np = ((const V4<T>*)D.u_pos)[slot];
    const V4<T> nv = ((const V4<T>*)D.u_vel)[slot];
    out[23] = (double)np.x; out[24] = (double)np.y; out[25] = (double)np.z;
    out[26] = (double)nv.x; out[27] = (double)nv.y; out[28] = (double)nv.z;
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// id -> sorted slot of the last step, for the scenes with a map (cleared to
// INT32_MAX first; duplicate ids keep their lowest slot).
__global__ __launch_bounds__(256) void k_idmap(DevState D, int32_t* __restrict__ map)
{
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= D.P) return;
    const SceneDev& S = D.scenes[wave_scene(D.chunk_scene, slot)];
    if (slot >= S.p_begin + S.n || S.idmap_off < 0) return;
    const int32_t id = D.s_id[slot];
    if (id >= 0 && id < S.idmap_len) atomicMin(&map[S.idmap_off + id], (int32_t)slot);
}

// Positions (what 0), velocities (1) or the last step's densities (2) of the
// n particles at slots p0 .. p0+n-1 in id order (the ids are 0 .. n-1): the
// current state lies at the last step's sorted slots with its ids in u_id.
// Each value is rounded to FP32 once, as the host copy rounds it.
template <typename T>
__global__ __launch_bounds__(256) void k_feed(DevState D, int64_t p0, int64_t n, int what, float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t slot = p0 + i;
    const int32_t id = D.u_id[slot];
    if (id < 0 || (int64_t)id >= n) return;
    if (what == 2) {
        out[id] = (float)((const T*)D.rho)[slot];
        return;
    }
    const V4<T> v = ((const V4<T>*)(what ? D.u_vel : D.u_pos))[slot];
    float* o = out + 3 * (int64_t)id;
    o[0] = (float)v.x;
    o[1] = (float)v.y;
    o[2] = (float)v.z;
}

// The sorted slot of particle `id` of one scene without a map (lowest slot
// for a duplicate id), into *slot_out (cleared to INT32_MAX first).
__global__ __launch_bounds__(256) void k_find_id(DevState D, int scene, int32_t id, int32_t* __restrict__ slot_out)
{
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= D.P) return;
    const int si = wave_scene(D.chunk_scene, slot);
    if (si != scene) return;
    const SceneDev& S = D.scenes[si];
    if (slot < S.p_begin + S.n && D.s_id[slot] == id) atomicMin(slot_out, (int32_t)slot);
}

// Acceleration of every particle of one scene in the current sorted state
// (density pass done): the force pass's sums (list, pool or walk, the same
// order and lane split) and finish(), without integration.
// out[3 (slot - p_begin) + a], sorted order.
template <typename T, int SM, int G>
__global__ __launch_bounds__(256) void k_accel(DevState D, int scene, double* __restrict__ out)
{
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t slot = gt / G;
    const int l = (int)(gt & (G - 1));
    if (slot >= D.P) return;
    const int si = wave_scene(D.chunk_scene, slot);
    if (si != scene) return;
    const SceneDev& S = D.scenes[si];
    if (slot >= S.p_begin + S.n) return;  // whole groups leave together
    const SceneConsts<T>& K = SceneSel<T>::get(S);
    const V4<T> pi = ((const V4<T>*)D.s_pos)[slot];
    const V4<T> vi = ((const V4<T>*)D.s_vel)[slot];
    const T rho_i = ((const T*)D.rho)[slot];
    const T a_i = ((const V4<T>*)D.posa)[slot].w;
    ForceAcc<T, SM> acc;
    if (l == 0) acc.self_term(K, vi.w);
    force_share<T, SM, G>(D, S, K, slot, l, pi, vi, a_i, acc, false);
    acc.template merge<G>();
    if (l != 0) return;
    ForceOut<T> o;
    acc.finish(K, rho_i, o);
#pragma unroll
    for (int a = 0; a < 3; ++a) out[3 * (slot - S.p_begin) + a] = (double)o.acc[a];
}

template <typename T>
__global__ __launch_bounds__(256) void k_neighbours(DevState D, int scene, const int64_t* __restrict__ off,
                                                    int32_t* __restrict__ counts, int32_t* __restrict__ idx)
{
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= D.P) return;
    const int si = D.chunk_scene[slot >> 6];
    if (si != scene) return;
    const SceneDev& S = D.scenes[si];
    if (slot >= S.p_begin + S.n) return;
    const SceneConsts<T>& K = SceneSel<T>::get(S);
    const V4<T>* pos = (const V4<T>*)D.s_pos;
    const V4<T> pi = pos[slot];
    const int64_t local = slot - S.p_begin;
    int32_t c = 0;
    const int64_t o = idx ? off[local] : 0;
    visit_neighbours<T>(D, S, K, pos, pi.x, pi.y, pi.z,
                        [&](uint32_t j, T, T, T, T) {
                            if (idx) idx[o + c] = D.s_id[j];
                            ++c;
                        });
    counts[local] = c;
}

template <typename T>
__global__ __launch_bounds__(256) void k_keys(DevState D, int scene, int32_t* __restrict__ keys)
{
    const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= D.P) return;
    const int si = D.chunk_scene[slot >> 6];
    if (si != scene) return;
    const SceneDev& S = D.scenes[si];
    if (slot >= S.p_begin + S.n) return;
    const SceneConsts<T>& K = SceneSel<T>::get(S);
    const V4<T> p = ((const V4<T>*)D.u_pos)[slot];
    int kx, ky, kz;
    cell_of<T>(K, S.dims, p.x, p.y, p.z, kx, ky, kz);
    keys[slot - S.p_begin] = (kz * S.dims[1] + ky) * S.dims[0] + kx;
}

// Grid::UpdateGrid's vertex stream (grid.h:138-176) of one scene's terrain,
// one thread per cell: the cell centre (org + (i + 0.5) cell), the height,
// and normalize(cross(u,l) + cross(u,r) + cross(d,l) + cross(d,r)) over the
// vectors to the edge neighbours present (normalize = v * (1 / |v|), as
// glm's; a zero sum gives a zero normal).
__global__ __launch_bounds__(256) void k_terrain_mesh(const float* __restrict__ H, int nx, int nz, float cell,
                                                      float ox, float oz, float* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * nz) return;
    const int x = (int)(t % nx), z = (int)(t / nx);
    const float y = H[t];
    float u[3] = {0.f, 0.f, 0.f}, d[3] = {0.f, 0.f, 0.f}, r[3] = {0.f, 0.f, 0.f}, l[3] = {0.f, 0.f, 0.f};
    if (x - 1 >= 0) { l[0] = cell; l[1] = y - H[t - 1]; }
    if (x + 1 < nx) { r[0] = cell; r[1] = H[t + 1] - y; }
    if (z - 1 >= 0) { u[1] = y - H[t - nx]; u[2] = cell; }
    if (z + 1 < nz) { d[1] = H[t + nx] - y; d[2] = cell; }
    const float* A[4] = {u, u, d, d};
    const float* B[4] = {l, r, l, r};
    float n[3];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float* a = A[k];
        const float* b = B[k];
        const float c0 = a[1] * b[2] - a[2] * b[1];
        const float c1 = a[2] * b[0] - a[0] * b[2];
        const float c2 = a[0] * b[1] - a[1] * b[0];
        if (k == 0) { n[0] = c0; n[1] = c1; n[2] = c2; }
        else { n[0] = n[0] + c0; n[1] = n[1] + c1; n[2] = n[2] + c2; }
    }
    const float len = d_sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
    const float inv = len > 0.f ? 1.f / len : 0.f;
    float* o = out + 6 * t;
    o[0] = ox + ((float)x + 0.5f) * cell;
    o[1] = y;
    o[2] = oz + ((float)z + 0.5f) * cell;
    o[3] = n[0] * inv;
    o[4] = n[1] * inv;
    o[5] = n[2] * inv;
}

__global__ void k_clear_status(uint32_t* status)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *status = 0u;
}

// Copies the list-overflow words (largest overflowing count, failed pool
// claims) and the pool entries the last density pass claimed into host memory
// mapped for the device (vector stores), where the host reads them between
// step sequences without waiting for the stream.
__global__ void k_publish(const uint32_t* __restrict__ ovf, const unsigned long long* __restrict__ ptop,
                          uint32_t* __restrict__ dst)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const unsigned long long t = *ptop;
    volatile uint32_t* d = dst;
    d[0] = ovf[0];
    d[1] = ovf[1];
    d[2] = (uint32_t)t;
    d[3] = (uint32_t)(t >> 32);
}

// Mass bookkeeping of one scene: FP64 sums of terrain volume and carried
// sediment, and the exact mod-2^64 sum of their fixed-point units.  Each
// block writes its partial; k_mass_final adds them in block order.
__global__ __launch_bounds__(256) void k_mass_partial(DevState D, int scene, double* __restrict__ part)
{
    __shared__ double s_t[256];
    __shared__ double s_s[256];
    __shared__ unsigned long long s_u[256];
    const SceneDev& S = D.scenes[scene];
    const int64_t ncells = (int64_t)S.tnx * S.tnz;
    const double unit = ldexp(1.0, -S.fx_exp);
    double st = 0.0, ss = 0.0;
    unsigned long long su = 0ull;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ncells; t += stride) {
        const int64_t v = D.vol[S.t_off + t];
        st += (double)v * unit;
        su += (unsigned long long)v;
    }
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < S.n; p += stride) {
        const int64_t v = D.u_sed[S.p_begin + p];
        ss += (double)v * unit;
        su += (unsigned long long)v;
    }
    s_t[threadIdx.x] = st;
    s_s[threadIdx.x] = ss;
    s_u[threadIdx.x] = su;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_t[threadIdx.x] += s_t[threadIdx.x + o];
            s_s[threadIdx.x] += s_s[threadIdx.x + o];
            s_u[threadIdx.x] += s_u[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = s_t[0];
        part[3 * blockIdx.x + 1] = s_s[0];
        part[3 * blockIdx.x + 2] = __longlong_as_double((long long)s_u[0]);
    }
}

// The same bookkeeping for every scene in one launch: block (b, s) takes
// cells and particles b, b + SPHE_MASS_BLOCKS_PER_SCENE blocks ... of scene
// s and writes its partial to part[3 (s B + b) ..]; k_mass_scenes_final adds
// a scene's B partials in block order.  Deterministic for a given layout.
// Every partial and result is written by each launch: the buffers need no
// clearing between calls.
__global__ __launch_bounds__(256) void k_mass_scenes(DevState D, double* __restrict__ part)
{
    __shared__ double s_t[256];
    __shared__ double s_s[256];
    __shared__ unsigned long long s_u[256];
    const int scene = (int)blockIdx.y;
    const SceneDev& S = D.scenes[scene];
    const int64_t ncells = (int64_t)S.tnx * S.tnz;
    const double unit = ldexp(1.0, -S.fx_exp);
    double st = 0.0, ss = 0.0;
    unsigned long long su = 0ull;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t first = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t t = first; t < ncells; t += stride) {
        const int64_t v = D.vol[S.t_off + t];
        st += (double)v * unit;
        su += (unsigned long long)v;
    }
    for (int64_t p = first; p < S.n; p += stride) {
        const int64_t v = D.u_sed[S.p_begin + p];
        ss += (double)v * unit;
        su += (unsigned long long)v;
    }
    s_t[threadIdx.x] = st;
    s_s[threadIdx.x] = ss;
    s_u[threadIdx.x] = su;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s_t[threadIdx.x] += s_t[threadIdx.x + o];
            s_s[threadIdx.x] += s_s[threadIdx.x + o];
            s_u[threadIdx.x] += s_u[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* q = part + 3 * ((int64_t)scene * gridDim.x + blockIdx.x);
        q[0] = s_t[0];
        q[1] = s_s[0];
        q[2] = __longlong_as_double((long long)s_u[0]);
    }
}

__global__ __launch_bounds__(256) void k_mass_scenes_final(const double* __restrict__ part, int nscenes, int nblocks,
                                                           double* __restrict__ out)
{
    const int s = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (s >= nscenes) return;
    double st = 0.0, ss = 0.0;
    unsigned long long su = 0ull;
    const double* q = part + 3 * (int64_t)s * nblocks;
    for (int b = 0; b < nblocks; ++b) {
        st += q[3 * b];
        ss += q[3 * b + 1];
        su += (unsigned long long)__double_as_longlong(q[3 * b + 2]);
    }
    // planar: terrain[nscenes], sediment[nscenes], units[nscenes]
    out[s] = st;
    out[nscenes + s] = ss;
    out[2 * nscenes + s] = __longlong_as_double((long long)su);
}

__global__ void k_mass_final(const double* __restrict__ part, int nblocks, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double st = 0.0, ss = 0.0;
    unsigned long long su = 0ull;
    for (int b = 0; b < nblocks; ++b) {
        st += part[3 * b];
        ss += part[3 * b + 1];
        su += (unsigned long long)__double_as_longlong(part[3 * b + 2]);
    }
    out[0] = st;
    out[1] = ss;
    out[2] = __longlong_as_double((long long)su);
}

inline unsigned grid_of(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace

// ------------------------------------------------------------- launchers

// The precision modes of sphe_create: 0 FP32 with Kahan-compensated sums,
// 1 FP32 plain, 2 FP64, 3 FP32 with compensated density sums and
// FP64-accumulated force sums.  KERN(T, SM) names the kernel instantiation.
#define SPHE_PREC_SWITCH(prec, CALL)                                      \
    do {                                                                  \
        if ((prec) == 0) { CALL(float, SUM_KAHAN); }                      \
        else if ((prec) == 1) { CALL(float, SUM_PLAIN); }                 \
        else if ((prec) == 3) { CALL(float, SUM_F64ACC); }                \
        else { CALL(double, SUM_PLAIN); }                                 \
    } while (0)

// Launches kern.  With profiling events, through hipExtLaunchKernel: e0 and
// e1 then take the kernel's own start and end times, with no marker packets
// between kernels.
template <typename... P, typename... A>
static void klaunch(void (*kern)(P...), dim3 grid, dim3 block, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                    A... args)
{
    if (e0 || e1) hipExtLaunchKernelGGL(kern, grid, block, 0, st, e0, e1, 0, args...);
    else hipLaunchKernelGGL(kern, grid, block, 0, st, args...);
}

// The same with lds bytes of dynamic LDS per block.
template <typename... P, typename... A>
static void klaunch_lds(void (*kern)(P...), dim3 grid, dim3 block, size_t lds, hipStream_t st, hipEvent_t e0,
                        hipEvent_t e1, A... args)
{
    if (e0 || e1) hipExtLaunchKernelGGL(kern, grid, block, lds, st, e0, e1, 0, args...);
    else hipLaunchKernelGGL(kern, grid, block, lds, st, args...);
}

// Blocks of `kern` (block threads, lds bytes of dynamic LDS) resident on
// the current device at once: the grid of a persistent launch (a multiple
// of the XCD count).
template <typename... P>
static int resident_grid(void (*kern)(P...), int block, size_t lds)
{
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, size_t, int>, int> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const auto key = std::make_tuple((const void*)kern, block, lds, dev);
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)kern, block, lds) != hipSuccess) per = 0;
    (void)hipGetLastError();
    const int n = std::max(8, (std::max(cus, 1) * std::max(per, 1)) / 8 * 8);
    cache[key] = n;
    return n;
}

// A pass as a persistent launch: the queues cleared, then a resident grid
// (at most one block per item).
template <typename... P, typename... A>
static void klaunch_pers(void (*kern)(P...), int nblk, int block, size_t lds, const DevState& D, hipStream_t st,
                         hipEvent_t e0, hipEvent_t e1, A... args)
{
    const int grid = std::min(resident_grid(kern, block, lds), nblk);
    klaunch(k_work_reset, 1, 64, st, e0, nullptr, D.work);
    klaunch_lds(kern, grid, block, lds, st, nullptr, e1, args...);
}

void sphe_launch_hash(const DevState& D, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    if (D.P == 0) return;
    if (D.precision == 2) klaunch(k_hash_count<double>, grid_of(D.P, 256), 256, st, e0, e1, D);
    else klaunch(k_hash_count<float>, grid_of(D.P, 256), 256, st, e0, e1, D);
}

template <int ITEMS>
static void launch_scan_t(uint32_t* cnt, uint32_t* start, int64_t n, uint32_t* partial, bool zero_counts,
                          unsigned long long* zero, unsigned long long zero_to, uint32_t* occ, uint32_t* fill,
                          hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    const int64_t nb = (n + 256 * ITEMS - 1) / (256 * ITEMS);
    klaunch(k_scan_reduce<ITEMS>, (unsigned)nb, 256, st, e0, nullptr, (const uint32_t*)cnt, n, partial,
            (const uint32_t*)occ);
    klaunch(k_scan_tiles, 1, 1024, st, nullptr, nullptr, partial, nb);
    klaunch(k_scan_down<ITEMS>, (unsigned)nb, 256, st, nullptr, e1, cnt, start, n, (const uint32_t*)partial,
            zero_counts ? 1 : 0, zero, zero_to, occ, occ ? fill : nullptr);
}

void sphe_launch_scan(uint32_t* cnt, uint32_t* start, int64_t n, uint32_t* partial, bool zero_counts,
                      unsigned long long* zero, unsigned long long zero_to, uint32_t* occ, uint32_t* fill,
                      int scan_log2, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    if (scan_log2 == 11) launch_scan_t<8>(cnt, start, n, partial, zero_counts, zero, zero_to, occ, fill, st, e0, e1);
    else if (scan_log2 == 13)
        launch_scan_t<32>(cnt, start, n, partial, zero_counts, zero, zero_to, occ, fill, st, e0, e1);
    else launch_scan_t<16>(cnt, start, n, partial, zero_counts, zero, zero_to, occ, fill, st, e0, e1);
}

int64_t sphe_scan_partials(int64_t n) { return (n + (1 << SPHE_SCAN_LOG2_MIN) - 1) >> SPHE_SCAN_LOG2_MIN; }

void sphe_launch_scatter(const DevState& D, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    if (D.P == 0) return;
    klaunch(k_scatter, grid_of(D.P, 256), 256, st, e0, e1, D);
}

void sphe_launch_reorder(const DevState& D, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    if (D.P == 0) return;
    if (D.precision == 2) klaunch(k_reorder<double>, grid_of(D.P, 256), 256, st, e0, e1, D);
    else klaunch(k_reorder<float>, grid_of(D.P, 256), 256, st, e0, e1, D);
}

template <int G>
static void launch_density_g(const DevState& D, int block, int remap, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    // one row table and one list staging ring per group, at most 64 KiB per
    // block with the persistent form's static word (halve the block until
    // it fits)
    const int W = 2 * D.yz + 1, ring = 4 * G > 16 ? 64 : 32;
    while (block > 64 && density_lds_bytes(W, block / G, ring) > 65536 - 256) block /= 2;
    const size_t lds = (size_t)density_lds_bytes(W, block / G, ring);
    const int nblk = (int)grid_of(D.P * G, block);
    if (D.persistent) {
#define CALL(T, SM) \
    klaunch_pers(k_density_pers<T, SM, G>, nblk, block, lds, D, st, e0, e1, D, nblk, remap, D.work)
        SPHE_PREC_SWITCH(D.precision, CALL);
#undef CALL
        return;
    }
#define CALL(T, SM) klaunch_lds(k_density<T, SM, G>, nblk, block, lds, st, e0, e1, D, nblk, remap)
    SPHE_PREC_SWITCH(D.precision, CALL);
#undef CALL
}

template <int G>
static void launch_force_g(const DevState& D, int block, int remap, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    const int nblk = (int)grid_of(D.P * G, block);
    if (D.persistent) {
        if (D.precision == 2) {
            klaunch_pers(k_force_f64_pers<G>, nblk, block, 0, D, st, e0, e1, D, nblk, remap, D.work);
            return;
        }
#define CALL(T, SM) klaunch_pers(k_force_f32_pers<SM, G>, nblk, block, 0, D, st, e0, e1, D, nblk, remap, D.work)
        SPHE_PREC_SWITCH(D.precision, CALL);
#undef CALL
        return;
    }
    if (D.precision == 2) {
        klaunch(k_force_f64<G>, nblk, block, st, e0, e1, D, nblk, remap);
        return;
    }
#define CALL(T, SM) klaunch(k_force_f32<SM, G>, nblk, block, st, e0, e1, D, nblk, remap)
    SPHE_PREC_SWITCH(D.precision, CALL);
#undef CALL
}

void sphe_launch_density(const DevState& D, int block, int remap, int lanes, hipStream_t st, hipEvent_t e0,
                         hipEvent_t e1)
{
    if (D.P == 0) return;
    switch (lanes) {
    case 1: launch_density_g<1>(D, block, remap, st, e0, e1); break;
    case 2: launch_density_g<2>(D, block, remap, st, e0, e1); break;
    case 8: launch_density_g<8>(D, block, remap, st, e0, e1); break;
    default: launch_density_g<4>(D, block, remap, st, e0, e1); break;
    }
}

void sphe_launch_force(const DevState& D, int block, int remap, int lanes, hipStream_t st, hipEvent_t e0,
                       hipEvent_t e1)
{
    if (D.P == 0) return;
    switch (lanes) {
    case 1: launch_force_g<1>(D, block, remap, st, e0, e1); break;
    case 2: launch_force_g<2>(D, block, remap, st, e0, e1); break;
    case 8: launch_force_g<8>(D, block, remap, st, e0, e1); break;
    default: launch_force_g<4>(D, block, remap, st, e0, e1); break;
    }
}

template <int ITEMS>
static void launch_scan_heights_t(const DevState& D, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    const int64_t n = D.C + 1;
    const int64_t nb = (n + 256 * ITEMS - 1) / (256 * ITEMS);
    const int64_t nh = (int64_t)grid_of(D.TC / 4, 256);  // TC: a multiple of 64
    klaunch(k_scan_reduce_heights<ITEMS>, (unsigned)(nb + nh), 256, st, e0, nullptr, D, (int)nb);
    klaunch(k_scan_tiles, 1, 1024, st, nullptr, nullptr, D.partial, nb);
    klaunch(k_scan_down<ITEMS>, (unsigned)nb, 256, st, nullptr, e1, D.cnt, D.start, n, (const uint32_t*)D.partial,
            0, D.ptop, (unsigned long long)D.pool_origin, D.tocc, D.tfill);
}

void sphe_launch_scan_heights(const DevState& D, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    if (D.scan_log2 == 11) launch_scan_heights_t<8>(D, st, e0, e1);
    else if (D.scan_log2 == 13) launch_scan_heights_t<32>(D, st, e0, e1);
    else launch_scan_heights_t<16>(D, st, e0, e1);
}

void sphe_launch_heights(const DevState& D, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    if (D.TC == 0) return;
    klaunch(k_heights, grid_of(D.TC / 4, 256), 256, st, e0, e1, D);  // TC: a multiple of 64
}

void sphe_launch_two_phase(const DevState& D, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    if (D.TC == 0 || D.P == 0) return;
    klaunch(k_tcount, grid_of(D.P, 256), 256, st, e0, nullptr, D);
    sphe_launch_scan(D.tcnt, D.tstart, D.TC + 1, D.tpartial, true, nullptr, 0ull, nullptr, nullptr,
                     SPHE_SCAN_LOG2_DEFAULT, st);
    k_tscatter<<<grid_of(D.P, 256), 256, 0, st>>>(D);
    klaunch(k_tapply, grid_of(D.TC, 256), 256, st, nullptr, e1, D);
}

void sphe_launch_clear_status(const DevState& D, hipStream_t st) { k_clear_status<<<1, 64, 0, st>>>(D.status); }

void sphe_launch_publish(const DevState& D, uint32_t* host_word, hipStream_t st)
{
    k_publish<<<1, 64, 0, st>>>(D.ovf, D.ptop, host_word);
}

void sphe_launch_terrain_mesh(const float* H, int nx, int nz, float cell, float ox, float oz, float* out,
                              hipStream_t st)
{
    const int64_t n = (int64_t)nx * nz;
    if (n == 0) return;
    k_terrain_mesh<<<grid_of(n, 256), 256, 0, st>>>(H, nx, nz, cell, ox, oz, out);
}

template <int G>
static void launch_inspect_g(const DevState& D, const int32_t* slotp, double* out, hipStream_t st)
{
#define CALL(T, SM) k_inspect<T, SM, G><<<1, 64, 0, st>>>(D, slotp, out)
    SPHE_PREC_SWITCH(D.precision, CALL);
#undef CALL
}

void sphe_launch_inspect(const DevState& D, const int32_t* slotp, int lanes, double* out, hipStream_t st)
{
    switch (lanes) {
    case 1: launch_inspect_g<1>(D, slotp, out, st); break;
    case 2: launch_inspect_g<2>(D, slotp, out, st); break;
    case 8: launch_inspect_g<8>(D, slotp, out, st); break;
    default: launch_inspect_g<4>(D, slotp, out, st); break;
    }
}

void sphe_launch_idmap(const DevState& D, int32_t* map, int64_t len, hipStream_t st)
{
    if (len > 0) k_fill_i32<<<grid_of(len, 256), 256, 0, st>>>(map, len, INT32_MAX);
    if (D.P > 0 && len > 0) k_idmap<<<grid_of(D.P, 256), 256, 0, st>>>(D, map);
}

void sphe_launch_feed(const DevState& D, int64_t p0, int64_t n, int what, float* out, hipStream_t st)
{
    if (n <= 0) return;
    if (D.precision == 2) k_feed<double><<<grid_of(n, 256), 256, 0, st>>>(D, p0, n, what, out);
    else k_feed<float><<<grid_of(n, 256), 256, 0, st>>>(D, p0, n, what, out);
}

void sphe_launch_find_id(const DevState& D, int scene, int32_t id, int32_t* slot_out, hipStream_t st)
{
    k_fill_i32<<<1, 64, 0, st>>>(slot_out, 1, INT32_MAX);
    if (D.P > 0) k_find_id<<<grid_of(D.P, 256), 256, 0, st>>>(D, scene, id, slot_out);
}

template <int G>
static void launch_accel_g(const DevState& D, int scene, double* out, hipStream_t st)
{
    const unsigned nb = grid_of(D.P * G, 256);
#define CALL(T, SM) k_accel<T, SM, G><<<nb, 256, 0, st>>>(D, scene, out)
    SPHE_PREC_SWITCH(D.precision, CALL);
#undef CALL
}

void sphe_launch_accel(const DevState& D, int scene, int lanes, double* out, hipStream_t st)
{
    if (D.P == 0) return;
    switch (lanes) {
    case 1: launch_accel_g<1>(D, scene, out, st); break;
    case 2: launch_accel_g<2>(D, scene, out, st); break;
    case 8: launch_accel_g<8>(D, scene, out, st); break;
    default: launch_accel_g<4>(D, scene, out, st); break;
    }
}

void sphe_launch_neighbours(const DevState& D, int scene, const int64_t* off, int32_t* counts, int32_t* idx,
                            hipStream_t st)
{
    if (D.P == 0) return;
    if (D.precision == 2) k_neighbours<double><<<grid_of(D.P, 256), 256, 0, st>>>(D, scene, off, counts, idx);
    else k_neighbours<float><<<grid_of(D.P, 256), 256, 0, st>>>(D, scene, off, counts, idx);
}

void sphe_launch_keys(const DevState& D, int scene, int32_t* keys, hipStream_t st)
{
    if (D.P == 0) return;
    if (D.precision == 2) k_keys<double><<<grid_of(D.P, 256), 256, 0, st>>>(D, scene, keys);
    else k_keys<float><<<grid_of(D.P, 256), 256, 0, st>>>(D, scene, keys);
}

int sphe_mass_blocks() { return 256; }

// Scratch of sphe_mass_all_noalloc, for callers that cannot allocate device
// memory: the per-scene block partials and the per-scene results of the
// all-scene bookkeeping for up to SPHE_NOALLOC_MAX_SCENES scenes.  Every GPU
// the library runs on holds its own copy until the process ends.  Host code
// uses the address sphe_noalloc_scratch returns, never this name.
static __device__ __align__(256) double scratch[(size_t)3 * (SPHE_MASS_BLOCKS_PER_SCENE + 1) *
                                                SPHE_NOALLOC_MAX_SCENES];

int sphe_noalloc_scratch(int device, double** ptr, size_t* doubles)
{
    static std::mutex mu;
    static std::map<int, double*> addr;  // device -> address of its scratch
    std::lock_guard<std::mutex> g(mu);
    auto it = addr.find(device);
    if (it == addr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            *ptr = nullptr;
            *doubles = 0;
            return -1;
        }
        it = addr.emplace(device, (double*)p).first;
    }
    *ptr = it->second;
    *doubles = sizeof(scratch) / sizeof(double);
    return 0;
}

void sphe_launch_mass(const DevState& D, int scene, double* part, double* out, hipStream_t st)
{
    k_mass_partial<<<sphe_mass_blocks(), 256, 0, st>>>(D, scene, part);
    k_mass_final<<<1, 64, 0, st>>>(part, sphe_mass_blocks(), out);
}

void sphe_launch_mass_scenes(const DevState& D, double* part, double* out, hipStream_t st)
{
    if (D.nscenes == 0) return;
    k_mass_scenes<<<dim3(SPHE_MASS_BLOCKS_PER_SCENE, (unsigned)D.nscenes), 256, 0, st>>>(D, part);
    k_mass_scenes_final<<<grid_of(D.nscenes, 256), 256, 0, st>>>(part, D.nscenes, SPHE_MASS_BLOCKS_PER_SCENE, out);
}

// ------------------------------------------------------------ calibration

namespace {
// Plain streaming copy (16 bytes per thread per iteration): the reference
// point for HBM bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE
// counters of the step's kernels.
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ a, float4* __restrict__ b, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}
}  // namespace

void sphe_launch_copy(const void* a, void* b, int64_t n16, hipStream_t st)
{
    k_copy<<<2048, 256, 0, st>>>((const float4*)a, (float4*)b, n16);
}
