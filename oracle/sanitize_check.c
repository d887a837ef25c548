/*
 * sanitize_check.c -- drives the oracle (test infrastructure only) through
 * its code paths for a build under AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_oracle_sanitized.py builds and runs
 * it): the reference restatement ref_step on the reference lattice, and the
 * grid step orc_step on a dam break over an eroding terrain in FP64, FP32
 * plain, FP32 compensated and FP32 with FP64 force sums, for every lane
 * count, slab count and y/z split.  Exits non-zero if the fixed-point water
 * and sediment units are not conserved or a step reports an error.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sph_oracle.h"

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;

static double uniform01(void)
{
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

static int check_reference(void)
{
    ref_params p;
    ref_default_params(&p);
    p.dt = 0.01f;
    const float origin[3] = {0.0f, 0.0f, 0.0f};
    ref_particle* P = (ref_particle*)calloc(1000, sizeof(ref_particle));
    if (!P) return 1;
    const int n = ref_lattice(1000, origin, 0, P, 1000);
    for (int s = 0; s < 2; ++s) ref_step(P, n, &p);
    int bad = 0;
    for (int i = 0; i < n; ++i)
        if (!isfinite(P[i].Density) || !(P[i].Density > 0.0f)) ++bad;
    free(P);
    return bad != 0;
}

/* A dam break of N particles falling onto a T x T eroding terrain, stepped
 * `steps` times; returns non-zero on an error, a change of the unit sum or
 * no sediment taken up at all. */
static int check_grid(int real, int comp, int lanes, int slabs, int yz, int steps)
{
    enum { N = 600, T = 12 };
    orc_scene s;
    memset(&s, 0, sizeof(s));
    s.h = 0.0457;
    s.mass = 0.02;
    s.rho0 = 998.29;
    s.k = 3.0;
    s.visc = 3.5;
    s.surf_tens = 0.0728;
    s.g[1] = -9.8;
    s.dt = 2e-3;
    s.box_c[0] = s.box_c[1] = s.box_c[2] = 0.2;
    s.box_e = 0.2;
    s.tnx = s.tnz = T;
    s.t_org[0] = 0.0;
    s.t_org[1] = 0.0;
    s.t_org[2] = 0.0;
    s.t_cell = 0.4 / T;
    s.fx_exp = 40;
    s.erosion = 1;
    s.kc = 0.05;
    s.ke = 0.3;
    s.kd = 0.3;
    s.band = 0.0457;
    s.lanes = lanes;
    s.slabs = slabs;
    s.yz = yz;
    double* pos = (double*)malloc(sizeof(double) * 3 * N);
    double* vel = (double*)calloc(3 * N, sizeof(double));
    int64_t* sed = (int64_t*)calloc(N, sizeof(int64_t));
    int32_t* ids = (int32_t*)malloc(sizeof(int32_t) * N);
    int64_t* vol = (int64_t*)malloc(sizeof(int64_t) * T * T);
    float* height = (float*)malloc(sizeof(float) * T * T);
    if (!pos || !vel || !sed || !ids || !vol || !height) return 1;
    const double sp = 0.02;
    for (int i = 0; i < N; ++i) {
        const int x = i % 10, z = (i / 10) % 10, y = i / 100;
        pos[3 * i] = 0.02 + x * sp + (uniform01() - 0.5) * 0.002;
        pos[3 * i + 1] = 0.08 + y * sp + (uniform01() - 0.5) * 0.002;
        pos[3 * i + 2] = 0.02 + z * sp + (uniform01() - 0.5) * 0.002;
        vel[3 * i + 1] = -0.5;
        ids[i] = i;
        if (real == 4)
            for (int a = 0; a < 3; ++a) pos[3 * i + a] = (double)(float)pos[3 * i + a];
    }
    orc_consts c;
    orc_derive(&s, &c);
    /* terrain column heights 0.03 .. 0.07 of the cell edge's area units */
    const double area = s.t_cell * s.t_cell;
    for (int t = 0; t < T * T; ++t) {
        const double hgt = 0.03 + 0.04 * uniform01();
        vol[t] = (int64_t)llrint(hgt * area * ldexp(1.0, s.fx_exp));
    }
    orc_heights(&c, T * T, vol, height);
    int64_t units0 = 0;
    for (int t = 0; t < T * T; ++t) units0 += vol[t];
    int rc = 0;
    for (int k = 0; k < steps && rc >= 0; ++k)
        rc = orc_step(&s, real, comp, N, pos, vel, sed, ids, vol, height, NULL);
    int64_t units1 = 0, carried = 0;
    for (int t = 0; t < T * T; ++t) units1 += vol[t];
    for (int i = 0; i < N; ++i) carried += sed[i];
    units1 += carried;
    free(pos);
    free(vel);
    free(sed);
    free(ids);
    free(vol);
    free(height);
    if (rc < 0) return 1;
    return units1 != units0 || carried == 0;  /* conserved, and the exchange did run */
}

int main(void)
{
    int failed = check_reference();
    if (failed) fprintf(stderr, "reference restatement: non-finite or non-positive density\n");
    const int reals[4][2] = {{8, 0}, {4, 0}, {4, 1}, {4, 2}};
    for (int m = 0; m < 4; ++m)
        for (int lanes = 1; lanes <= 8; lanes *= 2)
            for (int slabs = 1; slabs <= 4; slabs *= 2)
                for (int yz = 1; yz <= 3; ++yz) {
                    const int r = check_grid(reals[m][0], reals[m][1], lanes, slabs, yz, 3);
                    if (r) {
                        fprintf(stderr, "grid step real %d comp %d lanes %d slabs %d yz %d failed\n", reals[m][0],
                                reals[m][1], lanes, slabs, yz);
                        failed = 1;
                    }
                }
    if (!failed) printf("oracle paths clean\n");
    return failed;
}
