/* A C caller of the system C ABI through its static instance:
 * set_chimes_serial(small) -> init_chimes_serial(param, &rank) ->
 * calculate_chimes(...), on a configuration read from an .xyz file whose
 * comment line holds the cell vectors.  It compiles against
 * include/chimescalc_serial_C.h and links libchimescalc.so only.
 *
 *   serial_c_main <parameter file> <xyz file> <small 0/1>
 *
 * Prints the energy, the 9-component stress and the forces at full
 * precision. */
#include <stdio.h>
#include <stdlib.h>

#include "chimescalc_serial_C.h"

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s <parameter file> <xyz file> <small 0/1>\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[2], "r");
    if (!f) {
        fprintf(stderr, "cannot open %s\n", argv[2]);
        return 2;
    }
    int n = 0, rank = 0;
    double ca[3], cb[3], cc[3];
    char buf[4096];
    if (!fgets(buf, sizeof buf, f) || sscanf(buf, "%d", &n) != 1 || n <= 0) return 2;
    if (!fgets(buf, sizeof buf, f) ||
        sscanf(buf, "%lf %lf %lf %lf %lf %lf %lf %lf %lf", &ca[0], &ca[1], &ca[2], &cb[0], &cb[1], &cb[2], &cc[0],
               &cc[1], &cc[2]) != 9)
        return 2;
    double *x = calloc(n, sizeof(double)), *y = calloc(n, sizeof(double)), *z = calloc(n, sizeof(double));
    double *fx = calloc(n, sizeof(double)), *fy = calloc(n, sizeof(double)), *fz = calloc(n, sizeof(double));
    char **types = calloc(n, sizeof(char *));
    for (int i = 0; i < n; i++) {
        types[i] = calloc(64, 1);
        if (!fgets(buf, sizeof buf, f) || sscanf(buf, "%63s %lf %lf %lf", types[i], &x[i], &y[i], &z[i]) != 4)
            return 2;
    }
    fclose(f);
    double energy = 0.0, stress[9] = {0};
    set_chimes_serial(atoi(argv[3]));
    init_chimes_serial(argv[1], &rank);
    calculate_chimes(n, x, y, z, types, ca, cb, cc, &energy, fx, fy, fz, stress);
    printf("energy %.17g\n", energy);
    printf("stress");
    for (int k = 0; k < 9; k++) printf(" %.17g", stress[k]);
    printf("\n");
    for (int i = 0; i < n; i++) printf("force %.17g %.17g %.17g\n", fx[i], fy[i], fz[i]);
    return 0;
}
