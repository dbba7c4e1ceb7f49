/*
 * pt_capi_example.c -- a plain C caller of libpentrack_gpu.so through include/pentrack_gpu.h.
 *
 * What a C or C++ host that replaces TTracker::IntegrateParticle (include/tracking.h:46-47,
 * called from src/main.cpp:159) does with the library, with no Python involved: build the
 * world model (a 1 m cube of material "vacuum" inside the default solid "wall", the
 * reference's in/config.in layout of materials and solids), describe the species, generate
 * particles, track them with pt_track_batch and read back the end states and hit logs.
 *
 * usage: pt_capi_example N SEED OUT.bin
 *   Tracks N neutrons for 2 s and writes, little-endian:
 *     char magic[4] = "PTCX"; int32 n, n_tri, hit_cap, pad; uint64 seed; double tmax;
 *     float tri[n_tri][3][3]; double mats[2][9];
 *     double t0[n], y0[n][9], spin0[n][5]; uint64 pid[n];
 *     pt_particle_out out[n]; pt_hit hits[n][hit_cap]
 *   so that a test can rerun exactly these inputs through another path.
 * Exit status 0 on success; on a library error the message of pt_last_error() is printed.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/pentrack_gpu.h"

#define HIT_CAP 32

static uint64_t sm_state;

/* splitmix64: a small, self-contained generator for the example's particles */
static uint64_t splitmix64(void) {
    uint64_t z = (sm_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static double uniform(double lo, double hi) { return lo + (hi - lo) * (double)(splitmix64() >> 11) * 0x1p-53; }

static int check(int rc, const char *what) {
    if (rc != PT_OK) fprintf(stderr, "%s failed (%d): %s\n", what, rc, pt_last_error());
    return rc;
}

/* 12 triangles of the box [lo, hi], outward normals */
static void box_triangles(const double lo[3], const double hi[3], float *tri) {
    const double P[8][3] = {{lo[0], lo[1], lo[2]}, {hi[0], lo[1], lo[2]}, {hi[0], hi[1], lo[2]}, {lo[0], hi[1], lo[2]},
                            {lo[0], lo[1], hi[2]}, {hi[0], lo[1], hi[2]}, {hi[0], hi[1], hi[2]}, {lo[0], hi[1], hi[2]}};
    const int quads[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {2, 3, 7, 6}, {1, 2, 6, 5}, {0, 4, 7, 3}};
    int t = 0;
    for (int q = 0; q < 6; ++q) {
        const int v[2][3] = {{quads[q][0], quads[q][1], quads[q][2]}, {quads[q][0], quads[q][2], quads[q][3]}};
        for (int h = 0; h < 2; ++h, ++t)
            for (int c = 0; c < 3; ++c)
                for (int a = 0; a < 3; ++a) tri[9 * t + 3 * c + a] = (float)P[v[h][c]][a];
    }
}

int main(int argc, char **argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: %s N SEED OUT.bin\n", argv[0]);
        return 2;
    }
    const int n = atoi(argv[1]);
    const uint64_t seed = strtoull(argv[2], NULL, 10);
    if (n < 0) return 2;
    sm_state = seed;
    const double tmax = 2.0;

    /* world model: material table (include/geometry.h:20-31), default solid = wall */
    const int n_tri = 12;
    float tri[12 * 9];
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    box_triangles(lo, hi, tri);
    int32_t tri_solid[12] = {0};
    const uint32_t solid_id[1] = {2};
    const int32_t solid_mat[1] = {1};
    const int32_t ign_off[2] = {0, 0};
    /* FermiReal FermiImag DiffProb SpinflipProb RMSRoughness CorrelLength InternalBField
       ModifiedLambertProb LossPerBounce */
    const double mats[2][9] = {{335.0, 0, 0, 0, 0, 0, 0, 0, 0}, {0, 0, 0.1, 0, 0, 0, 0, 0, 0}};
    pt_geometry *geom = NULL;
    pt_fields *fields = NULL;
    if (check(pt_geometry_create(n_tri, tri, tri_solid, 1, solid_id, solid_mat, ign_off, NULL, 0, 2, &mats[0][0], 4,
                                 &geom), "pt_geometry_create"))
        return 1;
    if (check(pt_fields_create(0, NULL, 0, &fields), "pt_fields_create")) return 1;

    /* neutron constants of the reference (src/globals.cpp:13-32, src/neutron.cpp:20) */
    const double ele_e = 1.602176487E-19;
    pt_species sp;
    memset(&sp, 0, sizeof sp);
    sp.kind = 0;
    sp.q = 0.0;
    sp.m = 1.674927211E-27 / ele_e;
    sp.mu = -0.96623641e-26;
    sp.gamma = -1.83247185e8;
    sp.tau = 0.0;
    sp.tmax = 9e99;
    sp.lmax = 9e99;

    /* particles: uniform in the cube, kinetic energy uniform up to 300 neV, isotropic */
    int32_t *species_idx = calloc((size_t)n + 1, sizeof(int32_t));
    double *t0 = calloc((size_t)n + 1, sizeof(double)), *y0 = calloc((size_t)n * 9 + 1, sizeof(double));
    double *spin0 = calloc((size_t)n * 5 + 1, sizeof(double));
    uint64_t *pid = calloc((size_t)n + 1, sizeof(uint64_t));
    pt_particle_out *out = calloc((size_t)n + 1, sizeof(pt_particle_out));
    pt_hit *hits = calloc((size_t)n * HIT_CAP + 1, sizeof(pt_hit));
    if (!species_idx || !t0 || !y0 || !spin0 || !pid || !out || !hits) return 1;
    for (int i = 0; i < n; ++i) {
        double *y = y0 + 9 * i;
        for (int a = 0; a < 3; ++a) y[a] = uniform(0.01, 0.99);
        const double E = uniform(0.0, 300e-9);
        const double v = sqrt(2.0 * E / sp.m);     /* m in eV s^2 / m^2 */
        const double phi = uniform(0.0, 6.283185307179586), cth = uniform(-1.0, 1.0);
        const double sth = sqrt(1.0 - cth * cth);
        y[3] = v * sth * cos(phi);
        y[4] = v * sth * sin(phi);
        y[5] = v * cth;
        y[7] = (splitmix64() >> 63) ? 1.0 : -1.0;
        pid[i] = (uint64_t)i + 1;
    }

    if (check(pt_track_batch(NULL, geom, fields, 1, &sp, tmax, n, species_idx, t0, y0, spin0, seed, pid, out, hits,
                             HIT_CAP, NULL), "pt_track_batch"))
        return 1;

    int stopped = 0;
    long long nhit = 0;
    for (int i = 0; i < n; ++i) {
        stopped += out[i].stopID != PT_ID_UNKNOWN;
        nhit += out[i].Nhit;
    }
    printf("tracked %d neutrons: %d stopped, %lld wall hits\n", n, stopped, nhit);

    FILE *f = fopen(argv[3], "wb");
    if (!f) return 1;
    const int32_t hdr[4] = {n, n_tri, HIT_CAP, 0};
    fwrite("PTCX", 1, 4, f);
    fwrite(hdr, sizeof hdr, 1, f);
    fwrite(&seed, sizeof seed, 1, f);
    fwrite(&tmax, sizeof tmax, 1, f);
    fwrite(tri, sizeof(float), (size_t)n_tri * 9, f);
    fwrite(mats, sizeof mats, 1, f);
    fwrite(t0, sizeof(double), (size_t)n, f);
    fwrite(y0, sizeof(double), (size_t)n * 9, f);
    fwrite(spin0, sizeof(double), (size_t)n * 5, f);
    fwrite(pid, sizeof(uint64_t), (size_t)n, f);
    fwrite(out, sizeof(pt_particle_out), (size_t)n, f);
    fwrite(hits, sizeof(pt_hit), (size_t)n * HIT_CAP, f);
    if (fclose(f) != 0) return 1;

    pt_fields_destroy(fields);
    pt_geometry_destroy(geom);
    free(species_idx), free(t0), free(y0), free(spin0), free(pid), free(out), free(hits);
    return 0;
}
