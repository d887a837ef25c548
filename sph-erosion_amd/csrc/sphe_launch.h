// sphe_launch.h -- kernel launchers of sphe_kernels.hip used by the runtime.
#pragma once

#include <hip/hip_runtime.h>

#include "sphe_device.h"

// e0 / e1 (optional): profiling events that take the start of the first and
// the end of the last kernel a launcher issues.
void sphe_launch_hash(const DevState& D, hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// zero (optional): a word the scan sets to zero_to (the pool's claim counter);
// occ (optional): one flag per tile of 2^scan_log2 counts, set where counts were taken;
// fill (optional, with occ): per tile, the value an empty tile's starts already hold, or ~0
void sphe_launch_scan(uint32_t* cnt, uint32_t* start, int64_t n, uint32_t* partial, bool zero_counts,
                      unsigned long long* zero, unsigned long long zero_to, uint32_t* occ, uint32_t* fill,
                      int scan_log2, hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// tiles of n counts at the smallest tile size (the size of the per-tile arrays)
int64_t sphe_scan_partials(int64_t n);
// The particle scan with the terrain heights computed in its first launch.
void sphe_launch_scan_heights(const DevState& D, hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void sphe_launch_scatter(const DevState& D, hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void sphe_launch_reorder(const DevState& D, hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void sphe_launch_density(const DevState& D, int block, int remap, int lanes, hipStream_t st,
                         hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void sphe_launch_force(const DevState& D, int block, int remap, int lanes, hipStream_t st,
                       hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void sphe_launch_heights(const DevState& D, hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void sphe_launch_two_phase(const DevState& D, hipStream_t st, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
void sphe_launch_clear_status(const DevState& D, hipStream_t st);
// Copies D.ovf[0], D.ovf[1] and the 64-bit *D.ptop (low, high) to four host
// words mapped for the device.
void sphe_launch_publish(const DevState& D, uint32_t* host_word, hipStream_t st);
void sphe_launch_terrain_mesh(const float* H, int nx, int nz, float cell, float ox, float oz, float* out,
                              hipStream_t st);
// The 112-byte record's values of the particle at sorted slot *slotp (device
// memory) into out[0 .. 28]; out[22] = -1 when *slotp names no slot.
void sphe_launch_inspect(const DevState& D, const int32_t* slotp, int lanes, double* out, hipStream_t st);
// id -> sorted slot map of the last step for the scenes that have one.
void sphe_launch_idmap(const DevState& D, int32_t* map, int64_t len, hipStream_t st);
// Positions (what 0) or velocities (1) of the n particles at slots p0 ..
// p0+n-1, whose ids are 0 .. n-1, as 3 floats at out[3 id]; their last
// densities (2) as one float at out[id].
void sphe_launch_feed(const DevState& D, int64_t p0, int64_t n, int what, float* out, hipStream_t st);
// Sorted slot of one particle id of a scene by a search over its slots.
void sphe_launch_find_id(const DevState& D, int scene, int32_t id, int32_t* slot_out, hipStream_t st);
void sphe_launch_accel(const DevState& D, int scene, int lanes, double* out, hipStream_t st);
void sphe_launch_neighbours(const DevState& D, int scene, const int64_t* off, int32_t* counts, int32_t* idx,
                            hipStream_t st);
void sphe_launch_keys(const DevState& D, int scene, int32_t* keys, hipStream_t st);
int sphe_mass_blocks();
void sphe_launch_mass(const DevState& D, int scene, double* part, double* out, hipStream_t st);
// Blocks per scene of the all-scene bookkeeping: part holds
// 3 * SPHE_MASS_BLOCKS_PER_SCENE * nscenes doubles, out 3 * nscenes (terrain
// volumes of all scenes, then their sediment, then their unit sums).
#define SPHE_MASS_BLOCKS_PER_SCENE 32
void sphe_launch_mass_scenes(const DevState& D, double* part, double* out, hipStream_t st);
// Most scenes sphe_mass_all_noalloc serves from the library's static scratch
// (3 * (SPHE_MASS_BLOCKS_PER_SCENE + 1) doubles per scene, 3.2 MiB in all).
#define SPHE_NOALLOC_MAX_SCENES 4096
// Device address of that scratch on `device`, which must be the current
// device: looked up with hipGetSymbolAddress on the first call for the
// device and kept per device.  *doubles = its size.  0, or -1 when the
// lookup failed.
int sphe_noalloc_scratch(int device, double** ptr, size_t* doubles);
void sphe_launch_copy(const void* a, void* b, int64_t n16, hipStream_t st);
