/* fvc_zorder: renumbers an ompVC tetrahedral mesh and its velocity slices in
 * Z-order on the GPU (fvc_mesh_zorder, fvc_gather_rows).  Same command line
 * and output files as the reference tool (cudaVC/ZOrderSort/main.cpp:295-313):
 *
 *   fvc_zorder <coordinates.bin> <connectivity.bin> <adjacency.bin>
 *              <velocity prefix> <first> <last> <step>
 *
 * writes <file>.new beside each mesh file and <prefix>.<n>.bin.new for
 * n = first, first+step, ..., last.  The renumbered velocity files keep their
 * leading time stamp, so they stay readable by ompVC's slice loader
 * (velocity.c:184-211); the reference writes the stamp to its already-closed
 * input stream (main.cpp:287), so its velocity outputs lack it. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fvc.h"

static void die(const char *msg, const char *what) {
    fprintf(stderr, "Error: %s%s%s\n", msg, what ? " " : "", what ? what : "");
    exit(1);
}

/* int32 count, then count * per items of `size` bytes */
static void *read_counted(const char *path, size_t size, int per, int32_t *count) {
    FILE *f = fopen(path, "rb");
    if (!f) die("cannot open", path);
    if (fread(count, sizeof(int32_t), 1, f) != 1 || *count < 0) die("cannot read the count of", path);
    const size_t n = (size_t)*count * (size_t)per;
    void *buf = malloc(size * (n ? n : 1));
    if (!buf) die("out of memory reading", path);
    if (fread(buf, size, n, f) != n) die("short file", path);
    fclose(f);
    return buf;
}

static void write_counted(const char *path, int32_t count, const void *data, size_t size, int per) {
    char name[4096];
    snprintf(name, sizeof name, "%s.new", path);
    FILE *f = fopen(name, "wb");
    if (!f) die("cannot create", name);
    const size_t n = (size_t)count * (size_t)per;
    if (fwrite(&count, sizeof(int32_t), 1, f) != 1 || fwrite(data, size, n, f) != n) die("cannot write", name);
    fclose(f);
}

int main(int argc, char **argv) {
    if (argc != 8) {
        fprintf(stderr, "usage: %s coordinates.bin connectivity.bin adjacency.bin vel_prefix first last step\n",
                argv[0]);
        return 1;
    }
    const int first = atoi(argv[5]), last = atoi(argv[6]), step = atoi(argv[7]);
    if (step <= 0 && first <= last) die("velocity step must be positive", NULL);
    int32_t nn = 0, ne = 0, ne2 = 0;
    double *xyz = (double *)read_counted(argv[1], sizeof(double), 3, &nn);
    int32_t *conn = (int32_t *)read_counted(argv[2], sizeof(int32_t), 4, &ne);
    int32_t *adj = (int32_t *)read_counted(argv[3], sizeof(int32_t), 4, &ne2);
    if (ne2 != ne) die("numOfCells is not consistent", NULL);
    printf("numOfPoints = %d\nnumOfCells = %d\n", nn, ne);
    int32_t *norder = malloc(sizeof(int32_t) * (nn ? nn : 1)), *eorder = malloc(sizeof(int32_t) * (ne ? ne : 1));
    double *xyz_new = malloc(sizeof(double) * 3 * (nn ? nn : 1));
    int32_t *conn_new = malloc(sizeof(int32_t) * 4 * (ne ? ne : 1)), *adj_new = malloc(sizeof(int32_t) * 4 * (ne ? ne : 1));
    if (!norder || !eorder || !xyz_new || !conn_new || !adj_new) die("out of memory", NULL);
    int rc = fvc_mesh_zorder(nn, xyz, ne, conn, adj, norder, eorder, xyz_new, conn_new, adj_new);
    if (rc) die("renumbering failed:", fvc_strerror(rc));
    write_counted(argv[1], nn, xyz_new, sizeof(double), 3);
    write_counted(argv[2], ne, conn_new, sizeof(int32_t), 4);
    printf("Finish new connectivities\n");
    write_counted(argv[3], ne, adj_new, sizeof(int32_t), 4);
    printf("Finish new adjacencies\n");
    double *vin = malloc(sizeof(double) * 3 * (nn ? nn : 1)), *vout = malloc(sizeof(double) * 3 * (nn ? nn : 1));
    if (!vin || !vout) die("out of memory", NULL);
    for (int fr = first; fr <= last; fr += step) {
        char name[4096];
        snprintf(name, sizeof name, "%s.%d.bin", argv[4], fr);
        FILE *f = fopen(name, "rb");
        if (!f) die("cannot open", name);
        double stamp;
        if (fread(&stamp, sizeof(double), 1, f) != 1 || fread(vin, sizeof(double), 3 * (size_t)nn, f) != 3 * (size_t)nn)
            die("Fail to read velocities", name);
        fclose(f);
        if ((rc = fvc_gather_rows(nn, 3, norder, vin, vout))) die("velocity renumbering failed:", fvc_strerror(rc));
        snprintf(name, sizeof name, "%s.%d.bin.new", argv[4], fr);
        f = fopen(name, "wb");
        if (!f) die("cannot create", name);
        if (fwrite(&stamp, sizeof(double), 1, f) != 1 || fwrite(vout, sizeof(double), 3 * (size_t)nn, f) != 3 * (size_t)nn)
            die("cannot write", name);
        fclose(f);
    }
    printf("Finish new velocities\n");
    free(xyz);
    free(conn);
    free(adj);
    free(norder);
    free(eorder);
    free(xyz_new);
    free(conn_new);
    free(adj_new);
    free(vin);
    free(vout);
    return 0;
}
