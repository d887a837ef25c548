// ref_driver.cpp -- C entry points around the reference's own FluidSystemSPH.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference header Erosion/fluid_system.h where it lies in the reference
// checkout (nothing of it is copied here) into oracle/_ref/libsphref.so.  The
// tests use it to pin the oracle's restatement bit for bit, and
// tests/golden/make_golden.py uses it to write the committed fixtures.
//
// Build notes:
//  * No OpenMP: the reference's `collapse(2)` pragmas (fluid_system.h:108,
//    128, 159) are ill-formed for g++ and mean something else to clang; the
//    MSVC build ignores the clause (warning C4849), and every i is
//    independent, so the serial build gives the reference's results.
//  * <cmath> and `using std::abs` come first so that abs(float) in collisionS
//    (fluid_system.h:356-370) binds to the float overload, as with MSVC.
//  * Private members (h, k, len, smoothRadius, m_Particles) are reached
//    through the explicit-instantiation access rule, which is standard C++;
//    the header is used unmodified.
#include <cmath>
#include <cstdlib>
using std::abs;
#include "fluid_system.h"

#include <cstring>
#include <vector>

static_assert(sizeof(FluidParticle) == 112, "FluidParticle layout");

namespace {

template <typename Tag, typename Tag::type M>
struct Expose {
    friend typename Tag::type get(Tag) { return M; }
};

struct ParticlesTag { typedef std::vector<FluidParticle> FluidSystemSPH::*type; friend type get(ParticlesTag); };
struct HTag { typedef float FluidSystemSPH::*type; friend type get(HTag); };
struct KTag { typedef float FluidSystemSPH::*type; friend type get(KTag); };
struct LenTag { typedef float FluidSystemSPH::*type; friend type get(LenTag); };
struct RadiusTag { typedef float FluidSystemSPH::*type; friend type get(RadiusTag); };
struct IdTag { typedef int FluidSystemSPH::*type; friend type get(IdTag); };

template struct Expose<ParticlesTag, &FluidSystemSPH::m_Particles>;
template struct Expose<HTag, &FluidSystemSPH::h>;
template struct Expose<KTag, &FluidSystemSPH::k>;
template struct Expose<LenTag, &FluidSystemSPH::len>;
template struct Expose<RadiusTag, &FluidSystemSPH::smoothRadius>;
template struct Expose<IdTag, &FluidSystemSPH::id>;

// One long-lived instance: FluidSystemSPH's destructor would release GL
// objects of the render helpers, so it is never destroyed.
FluidSystemSPH& sys()
{
    static FluidSystemSPH* s = new FluidSystemSPH();
    return *s;
}

Grid& grid()
{
    static Grid* g = new Grid(1, 1, 1);
    return *g;
}

// prm = {h, mass, p0, k, visc, surf_tens, gx, gy, gz, dt, len}
void apply_params(FluidSystemSPH& f, const float* prm)
{
    f.*get(HTag()) = prm[0];
    f.*get(RadiusTag()) = prm[0];
    *f.GetMass() = prm[1];
    *f.Getp0() = prm[2];
    f.*get(KTag()) = prm[3];
    *f.GetVisc() = prm[4];
    *f.GetSurfTen() = prm[5];
    *f.GetGrav() = glm::vec3(prm[6], prm[7], prm[8]);
    f.SetDeltaTime(prm[9]);
    f.*get(LenTag()) = prm[10];
}

}  // namespace

extern "C" {

// Initialize(n) at origin, then AddParticles(add_n) if add_n > 0
// (fluid_system.h:74-102, 232-251).  Copies up to cap records to out and
// returns the record count.
int refx_lattice(int n, int add_n, const float* origin, void* out, int cap)
{
    FluidSystemSPH& f = sys();
    (f.*get(ParticlesTag())).clear();
    f.*get(IdTag()) = 0;
    f.SetOrigin(glm::vec3(origin[0], origin[1], origin[2]));
    f.Initialize(n);
    if (add_n > 0) f.AddParticles(add_n);
    std::vector<FluidParticle>& P = f.*get(ParticlesTag());
    int count = (int)P.size();
    if (out && cap >= count) std::memcpy(out, P.data(), sizeof(FluidParticle) * count);
    return count;
}

// Loads n records, applies prm, runs `steps` FluidSystemSPH::Run calls
// (fluid_system.h:104-183) and copies the records back.
void refx_step(void* recs, int n, const float* prm, int steps)
{
    FluidSystemSPH& f = sys();
    apply_params(f, prm);
    std::vector<FluidParticle>& P = f.*get(ParticlesTag());
    P.resize(n);
    std::memcpy(P.data(), recs, sizeof(FluidParticle) * n);
    for (int s = 0; s < steps; s++) f.Run(grid());
    std::memcpy(recs, P.data(), sizeof(FluidParticle) * n);
}

// The reference's terrain mesh: a Grid(d, d, d) loads the 512 x 512 uint8
// heightfield img (Grid::LoadHeightfield, grid.h:98-102) and runs
// Grid::UpdateGrid(d, d, d) (grid.h:138-176, with genIndices :118-136).
// Copies up to vcap floats of GetSurfaceParts and icap indices of
// GetIndices (grid.h:807-830); returns the index count, *nv the float count.
int refx_grid_mesh(const unsigned char* img, int d, float* verts, int vcap, unsigned int* idx, int icap, int* nv)
{
    Grid* g = new Grid(d, d, d);
    g->LoadHeightfield(const_cast<unsigned char*>(img));
    g->UpdateGrid(d, d, d);
    const std::vector<float> v = g->GetSurfaceParts();
    const std::vector<unsigned int> i = g->GetIndices();
    *nv = (int)v.size();
    if (verts && vcap >= (int)v.size()) std::memcpy(verts, v.data(), v.size() * sizeof(float));
    if (idx && icap >= (int)i.size()) std::memcpy(idx, i.data(), i.size() * sizeof(unsigned int));
    // Grid has no destructor (grid.h: raw arrays are never freed); the test
    // process is short-lived
    return (int)i.size();
}

}  // extern "C"
