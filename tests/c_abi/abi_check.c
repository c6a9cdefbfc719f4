/* Compiled as C99 against include/vtx.h and linked with velvet_amd/libvtx.so by
 * tests/test_c_abi.py. Uses only host-side entry points (no GPU): prints the struct sizes the
 * C compiler gives, the reference defaults, and the constraints of a folded 3 x 3 grid from
 * the mesh generator, one value per line, for the test to compare with the ctypes binding. */
#include <stdio.h>
#include <string.h>

#include "vtx.h"

int main(void) {
    printf("sizeof vtx_params %zu\n", sizeof(vtx_params));
    printf("sizeof vtx_collider %zu\n", sizeof(vtx_collider));
    printf("sizeof vtx_tuning %zu\n", sizeof(vtx_tuning));
    printf("sizeof vtx_server_config %zu\n", sizeof(vtx_server_config));
    printf("sizeof vtx_server_stats %zu\n", sizeof(vtx_server_stats));
    vtx_params p;
    memset(&p, 0, sizeof p);
    vtx_params_default(&p);
    printf("defaults %d %d %d %.3f\n", p.num_substeps, p.num_iterations, p.max_num_neighbors, p.gravity[1]);

    /* 3 x 3 vertex grid in the xz plane, the last column lifted so the middle edges fold */
    float xyz[27];
    uint32_t tris[24], t = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            float* v = xyz + 3 * (a * 3 + b);
            v[0] = (float)b;
            v[1] = b == 2 ? 0.5f : 0.0f;
            v[2] = (float)a;
        }
    for (uint32_t a = 0; a < 2; a++)
        for (uint32_t b = 0; b < 2; b++) {
            const uint32_t i = a * 3 + b;
            tris[t++] = i; tris[t++] = i + 1; tris[t++] = i + 3;
            tris[t++] = i + 3; tris[t++] = i + 1; tris[t++] = i + 4;
        }
    int32_t ij[64];
    float rest[32], brest[32];
    uint32_t bidx[128], ns = 0, nb = 0;
    int rc = vtx_mesh_constraints(xyz, 9, tris, 8, ij, rest, 32, &ns, bidx, brest, 32, &nb);
    printf("rc %d\n", rc);
    printf("counts %u %u\n", ns, nb);
    for (uint32_t k = 0; k < ns; k++) printf("stretch %d %d %.9g\n", ij[2 * k], ij[2 * k + 1], rest[k]);
    for (uint32_t k = 0; k < nb; k++)
        printf("bend %u %u %u %u %.9g\n", bidx[4 * k], bidx[4 * k + 1], bidx[4 * k + 2], bidx[4 * k + 3], brest[k]);
    /* a capacity too small is an error with a message, not a crash */
    rc = vtx_mesh_constraints(xyz, 9, tris, 8, ij, rest, 2, &ns, bidx, brest, 32, &nb);
    printf("small rc %d %s\n", rc, rc ? "error" : "ok");
    printf("message %s\n", vtx_last_error());
    printf("version %s\n", vtx_version());
    printf("abi %d %d\n", vtx_abi_version(), VTX_ABI_VERSION);
    return 0;
}
