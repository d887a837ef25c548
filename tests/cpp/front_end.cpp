// A front end written against the reference's FluidSystemSPH interface
// (the calls Erosion/main.cpp makes: SetOrigin, Initialize, SetDeltaTime,
// the parameter pointer getters, Run(grid), GetParticle, AddParticles,
// Reset), compiled against the drop-in header instead of fluid_system.h.
// Prints records the GPU test compares with the oracle.
#include <cstdio>
#include <cstdlib>

#include "sphe_fluid_system.hpp"

// The reference's Grid (grid.h) is the front end's terrain renderer; the
// step ignores it (fluid_system.h:104), so a stand-in type is enough here.
struct Grid {
    Grid(int, int, int) {}
};

static void print_particle(const FluidParticle& fp)
{
    std::printf("%d %.9g %.9g %.9g %.9g %.9g %.9g %.9g %d\n", fp.Id, fp.Position[0], fp.Position[1],
                fp.Position[2], fp.Velocity[0], fp.Velocity[1], fp.Velocity[2], fp.Density, fp.NeighbId);
}

int main(int argc, char** argv)
{
    const int steps = argc > 1 ? std::atoi(argv[1]) : 1;
    try {
        FluidSystemSPH fluidsph;
        const float origin[3] = {0.0f, 0.0f, 0.0f};
        fluidsph.SetOrigin(origin);  // main.cpp:170
        Grid grid(50, 255, 50);      // main.cpp:101
        fluidsph.Initialize(1000);   // main.cpp:183
        fluidsph.SetDeltaTime(0.01f);  // main.cpp:481
        *fluidsph.GetMass() = 0.02f;   // main.cpp:278-290: UI edits through the pointers
        *fluidsph.GetVisc() = 3.5f;
        *fluidsph.GetSurfTen() = 0.0728f;
        *fluidsph.Getp0() = 998.29f;
        float* g = fluidsph.GetGrav();
        g[0] = 0.0f;
        g[1] = -9.82f;
        g[2] = 0.0f;
        for (int i = 0; i < steps; ++i) fluidsph.Run(grid);  // main.cpp:313
        print_particle(fluidsph.GetParticle(555));            // main.cpp:293
        print_particle(fluidsph.GetParticle(0));
        fluidsph.AddParticles(125);  // main.cpp:507
        fluidsph.Run(grid);
        std::printf("n %d\n", fluidsph.NumParticles());
        fluidsph.Reset();  // main.cpp:501
        std::printf("n %d\n", fluidsph.NumParticles());
        bool threw = false;
        try {
            fluidsph.GetParticle(123456);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        std::printf("range_error %d\n", threw ? 1 : 0);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
