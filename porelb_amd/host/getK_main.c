/*
 * getK.x -- analytic permeability of a body-centred cubic array of spheres
 * (the reference's getK.c:39-95, which does not compile as shipped: the
 * usage printf lines at getK.c:53-54 lack semicolons).
 *
 *     ./getK.x Chi nx          prints K1=... and K2=...
 *     ./getK.x Chi nx table    also prints the reference's TABLE block
 *
 * With L = nx - 1 and a = sqrt(3) L Chi / 4 the sphere radius:
 *   porosity  eps = 1 - (8/3) pi a^3 / L^3   (two spheres per cell)
 *   surface   S   = 2 * 4 pi a^2 / L^3
 *   drag      d*  = sum_{i<31} qs_i Chi^i    (Sangani & Acrivos 1982, BCC)
 *   K1 = 2 a^2 / (9 d* (1 - eps))            Sangani-Acrivos
 *   K2 = eps^3 / (C S^2), C = 5              Kozeny-Carman
 * The TABLE block (getK.c:74-92) evaluates K1 with an extra factor eps, as
 * the reference does.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const double PI = 3.14159265358979323846;

/* Sangani & Acrivos (1982) dimensionless drag series for the BCC array,
 * coefficients of Chi^0 .. Chi^30 (the values tabulated in getK.c:5-37). */
static const double qs_bcc[31] = {
    1.0,      1.575834, 2.483254, 3.233022, 4.022864, 4.650320, 5.281412, 5.826374,
    6.258376, 6.544504, 6.878396, 7.190839, 7.268068, 7.304025, 7.301217, 7.236410,
    7.298014, 7.369849, 7.109497, 6.228418, 5.235796, 4.476874, 3.541982, 2.939353,
    3.935484, 5.179097, 3.959872, 2.227627, 3.393390, 4.491369, 2.200686};

static double drag(double chi)
{
    double d = 0.0;
    for (int i = 0; i < 31; i++) d += qs_bcc[i] * pow(chi, i);
    return d;
}

int main(int argc, char *argv[])
{
    if (argc != 3 && !(argc == 4 && strcmp(argv[3], "table") == 0)) {
        fprintf(stderr, "argc err\n");
        printf("Usage: ./getK.x Chi nx [table]\n");
        printf("0<Chi<1 controls the size of the sphere, and nx is the resolution\n");
        exit(-1);
    }
    const double chi = atof(argv[1]);
    const int nx = atoi(argv[2]);
    const double L = (double)(nx - 1), C = 5.0;
    double a = sqrt(3.0) * L * chi / 4.0;
    double eps = 1.0 - 8.0 / 3 * PI * a * a * a / L / L / L;
    double S = 2 * 4 * PI * a * a / L / L / L;
    printf("K1=%e\n", 2.0 / 9 * a * a / drag(chi) / (1 - eps));
    printf("K2=%e\n", eps * eps * eps / C / S / S);
    if (argc == 4) {
        printf("Chi\tepsilon\tK1\tK2\n");
        for (int r = 0; r < 31; r++) {
            const double c = r * (1.0 / 30);
            a = sqrt(3.0) * L * c / 4.0;
            S = 2 * 4 * PI * a * a / L / L / L;
            eps = 1.0 - 8.0 / 3 * PI * a * a * a / L / L / L;
            printf("%e\t%e\t%e\t%e\n", c, eps, 2.0 / 9 * a * a * eps / drag(c) / (1 - eps),
                   eps * eps * eps / C / S / S);
        }
    }
    return 0;
}
