/* A C program written against the reference's public C API for its Chebyshev
 * kernel (include/scalfmm_capi.h): uniform particles, one evaluation, the
 * energy and the first potential printed.  Usage: capi_demo [n [output]];
 * with an output path it writes n (int64), the positions (3 n, xyz per
 * particle), the charges, the potentials and the forces (3 n) as doubles.
 * Compiled and linked by tests/test_capi_build.py, run by
 * tests/test_gpu_integration.py. */
#define _XOPEN_SOURCE 600
#include <stdio.h>
#include <stdlib.h>

#include "scalfmm_capi.h"

int main(int argc, char** argv) {
    const int n = argc > 1 ? atoi(argv[1]) : 10000;
    double* xyz = malloc(sizeof(double) * 3 * n);
    double* q = malloc(sizeof(double) * n);
    double* pot = calloc(n, sizeof(double));
    double* f = calloc(3 * n, sizeof(double));
    srand48(1);
    for (int i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) xyz[3 * i + d] = drand48();
        q[i] = 2 * drand48() - 1;
    }
    double center[3] = {0.5, 0.5, 0.5};
    Scalfmm_Cell_Descriptor cell = {0};
    Scalfmm_Leaf_Descriptor leaf = {0};
    scalfmm_handle h = scalfmm_init(chebyshev, multi_thread);
    if (!h) return 1;
    scalfmm_build_tree(h, 5, 1.0, center, cell, leaf);
    scalfmm_tree_insert_particles_xyz(h, n, xyz, BOTH);
    scalfmm_set_physical_values(h, n, q, BOTH);
    scalfmm_execute_fmm(h);
    scalfmm_get_potentials(h, n, pot, BOTH);
    scalfmm_get_forces_xyz(h, n, f, BOTH);
    double energy = 0;
    for (int i = 0; i < n; ++i) energy += q[i] * pot[i];
    double timers[9] = {0};
    scalfmm_get_timers(h, timers);
    printf("n %d energy %.12e pot[0] %.12e fx[0] %.12e m2l %.3e s\n", n, energy, pot[0], f[0], timers[2]);
    if (argc > 2) {
        FILE* out = fopen(argv[2], "wb");
        if (!out) return 1;
        const long long nn = n;
        fwrite(&nn, sizeof(nn), 1, out);
        fwrite(xyz, sizeof(double), 3 * (size_t)n, out);
        fwrite(q, sizeof(double), (size_t)n, out);
        fwrite(pot, sizeof(double), (size_t)n, out);
        fwrite(f, sizeof(double), 3 * (size_t)n, out);
        fclose(out);
    }
    scalfmm_dealloc_handle(h, NULL);
    free(xyz);
    free(q);
    free(pot);
    free(f);
    return 0;
}
