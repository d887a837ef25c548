// sphe_fluid_system.hpp -- C++ drop-in for the reference's FluidSystemSPH.
//
// Same method names and signatures as Erosion/fluid_system.h:66-485, backed
// by the C ABI (sphe.h) and the MI355X kernels.  A front end written against
// the reference (Erosion/main.cpp) compiles against this header after
// replacing #include "fluid_system.h".  Two calls change: Draw (GL
// rendering, which stays with the front end: it draws the positions that
// GetVertices returns) and the optional terrain, which the reference does
// not couple to the fluid.
//
// Errors: a failing call throws std::runtime_error with sphe_last_error().
// When glm is included first (GLM_VERSION defined), GetGrav / GetOrigin /
// SetOrigin use glm::vec3 exactly like the reference.
#pragma once

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "sphe.h"

// FluidParticle (fluid_system.h:49-64): same 112-byte layout.
typedef sphe_particle FluidParticle;

class FluidSystemSPH {
public:
    // fluid_system.h:69-72
    explicit FluidSystemSPH(int precision = SPHE_FP32)
    {
        check(sphe_create(precision, &ctx_), "sphe_create");
        sphe_params p;
        sphe_params_default(&p);
        check(sphe_add_scene(ctx_, &p, nullptr, &scene_), "sphe_add_scene");
    }
    ~FluidSystemSPH()
    {
        if (ctx_) sphe_destroy(ctx_);
    }
    FluidSystemSPH(const FluidSystemSPH&) = delete;
    FluidSystemSPH& operator=(const FluidSystemSPH&) = delete;

    // fluid_system.h:74-102
    void Initialize(int nParts) { check(sphe_init_lattice(ctx_, scene_, nParts), "Initialize"); }

    // fluid_system.h:104-183.  The reference ignores its Grid argument; so
    // does this (terrain coupling is attached with LoadHeightfield).
    template <class GridT>
    void Run(GridT&)
    {
        Run();
    }
    void Run() { check(sphe_step(ctx_, 1), "Run"); }
    // several steps in one call (one graph launch per step, no host sync)
    void Run(int steps) { check(sphe_step(ctx_, steps), "Run"); }

    // fluid_system.h:206-214
#if defined(GLM_VERSION)
    void SetOrigin(const glm::vec3& o)
    {
        const float v[3] = {o.x, o.y, o.z};
        check(sphe_set_origin(ctx_, scene_, v), "SetOrigin");
    }
    glm::vec3 GetOrigin() const
    {
        float v[3];
        sphe_get_origin(ctx_, scene_, v);
        return glm::vec3(v[0], v[1], v[2]);
    }
#endif
    void SetOrigin(const float o[3]) { check(sphe_set_origin(ctx_, scene_, o), "SetOrigin"); }

    // fluid_system.h:216-224
    void SetDeltaTime(float dt) { check(sphe_set_dt(ctx_, scene_, dt), "SetDeltaTime"); }
    float GetDeltaTime() const
    {
        float dt = 0.0f;
        sphe_get_dt(ctx_, scene_, &dt);
        return dt;
    }

    // fluid_system.h:226-230
    void PrintCoords()
    {
        std::vector<float> p = GetVertices();
        for (size_t i = 0; i < p.size() / 3; ++i)
            std::printf("[%zu] %g %g %g\n", i, p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    }

    // fluid_system.h:232-259
    void AddParticles(int n) { check(sphe_add_particles(ctx_, scene_, n), "AddParticles"); }
    void Reset() { check(sphe_reset(ctx_, scene_), "Reset"); }

    // fluid_system.h:261-284: writable parameter storage read at every step
    float* GetMass() { return sphe_param_ptr(ctx_, scene_, SPHE_P_MASS); }
    float* GetVisc() { return sphe_param_ptr(ctx_, scene_, SPHE_P_VISC); }
    float* GetSurfTen() { return sphe_param_ptr(ctx_, scene_, SPHE_P_SURF_TEN); }
    float* Getp0() { return sphe_param_ptr(ctx_, scene_, SPHE_P_P0); }
#if defined(GLM_VERSION)
    glm::vec3* GetGrav() { return reinterpret_cast<glm::vec3*>(sphe_param_ptr(ctx_, scene_, SPHE_P_GRAV)); }
#else
    float* GetGrav() { return sphe_param_ptr(ctx_, scene_, SPHE_P_GRAV); }
#endif

    // fluid_system.h:286-289
    FluidParticle GetParticle(int id)
    {
        FluidParticle p;
        check(sphe_get_particle(ctx_, scene_, id, &p), "GetParticle");
        return p;
    }

    // ---- additions beyond the reference's interface -------------------------

    // Positions, 3 floats per particle in id order (fluid_system.h:293-303,
    // getVertices): the buffer for one instanced draw in place of Draw.
    std::vector<float> GetVertices()
    {
        int n = 0;
        sphe_num_particles(ctx_, scene_, &n);
        std::vector<float> v((size_t)3 * n);
        check(sphe_read_buffer(ctx_, scene_, SPHE_BUF_POS, v.data(), v.size() * sizeof(float)), "GetVertices");
        return v;
    }
    int NumParticles() const
    {
        int n = 0;
        sphe_num_particles(ctx_, scene_, &n);
        return n;
    }

    // Terrain coupling (heightfield collision + erosion): Grid::LoadHeightfield
    // semantics (grid.h:98-107) with a height scale, then the eroded surface
    // in Grid::UpdateGrid / GetSurfaceParts / GetIndices form.  img, w, h as
    // stbi_load returns them (row-major, w pixels per row, h rows); cell
    // (x, z) is GetHeightfieldAt(x, z) = img[w * x + z]: h cells along x, w
    // along z.
    void LoadHeightfield(const unsigned char* img, int w, int h, float scale, float cell, const float origin[3],
                         int erosion = SPHE_EROSION_ATOMIC)
    {
        sphe_params p;
        check(sphe_get_params(ctx_, scene_, &p), "get_params");
        p.erosion = erosion;
        check(sphe_set_params(ctx_, scene_, &p), "set_params");
        sphe_terrain_desc t{};
        t.nx = h;
        t.nz = w;
        for (int a = 0; a < 3; ++a) t.origin[a] = origin[a];
        t.cell = cell;
        t.heights = nullptr;
        check(sphe_set_terrain(ctx_, scene_, &t), "LoadHeightfield");
        check(sphe_terrain_load_u8(ctx_, scene_, img, w, h, scale), "LoadHeightfield");
    }
    std::vector<float> GetSurfaceParts()
    {
        size_t nv = 0, ni = 0;
        check(sphe_terrain_mesh(ctx_, scene_, nullptr, &nv, nullptr, &ni), "GetSurfaceParts");
        std::vector<float> v(nv);
        check(sphe_terrain_mesh(ctx_, scene_, v.data(), &nv, nullptr, &ni), "GetSurfaceParts");
        return v;
    }
    std::vector<unsigned int> GetIndices()
    {
        size_t nv = 0, ni = 0;
        check(sphe_terrain_mesh(ctx_, scene_, nullptr, &nv, nullptr, &ni), "GetIndices");
        std::vector<unsigned int> idx(ni);
        check(sphe_terrain_mesh(ctx_, scene_, nullptr, &nv, idx.data(), &ni), "GetIndices");
        return idx;
    }

    sphe_ctx* context() { return ctx_; }
    int scene() const { return scene_; }

private:
    void check(int rc, const char* what)
    {
        if (rc != SPHE_OK) throw std::runtime_error(std::string(what) + ": " + sphe_last_error(ctx_));
    }

    sphe_ctx* ctx_ = nullptr;
    int scene_ = 0;
};
