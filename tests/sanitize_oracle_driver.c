/* sanitize_oracle_driver.c -- exercises the CPU oracle's field operations
 * (boundary resolution, ghost fills, steps, moments, initialisation) on small
 * boxes of every lattice and boundary kind, for a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_sanitize_cpu.py).
 * Test infrastructure: returns 0 when every call completed. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "../oracle/lbm_oracle.h"

static int run_case(int lat, const int n[3], const int face[6], const double fu[18], const double* F)
{
    const orc_lattice* L = orc_get_lattice(lat);
    const int D = L->D;
    int e[3];
    for (int d = 0; d < 3; ++d) e[d] = d < D ? n[d] + 2 : n[d];
    const size_t cells = (size_t)e[0] * e[1] * e[2], vcells = (size_t)n[0] * n[1] * n[2];
    double* fa = malloc(sizeof(double) * L->Q * cells);
    double* fb = malloc(sizeof(double) * L->Q * cells);
    double* rho = malloc(sizeof(double) * vcells);
    double* vel = malloc(sizeof(double) * D * vcells);
    int rc = 0;
    if (!fa || !fb || !rho || !vel) {
        rc = 1;
        goto done;
    }
    for (size_t i = 0; i < vcells; ++i) rho[i] = 1.0 + 0.01 * (double)(i % 7);
    for (size_t i = 0; i < D * vcells; ++i) vel[i] = 0.01 * (double)((int)(i % 5) - 2);
    for (size_t i = 0; i < L->Q * cells; ++i) fa[i] = fb[i] = NAN;
    orc_init_equilibrium(lat, rho, vel, fa, n, F);
    const double frho[6] = {1, 1.02, 1, 1, 1, 1};
    int rtype[27];
    double rrho[27], ru[81];
    int per[3] = {0, 0, 0};
    for (int d = 0; d < 3; ++d) per[d] = d < D ? face[2 * d] == 5 : 1;
    if (orc_resolve_bc(D, face, frho, fu, rtype, rrho, ru) != 0) {
        rc = 2;
        goto done;
    }
    /* ghosts of the initial state, then the advance() loop */
    orc_fill_periodic(lat, fa, n, per);
    if (orc_apply_bc(lat, fa, n, per, rtype, rrho, ru) < 0) {
        rc = 3;
        goto done;
    }
    const int last = orc_run(lat, fa, fb, n, 1.0 / 0.8, F, per, rtype, rrho, ru, 4, 2);
    if (last < 0) {
        rc = 4;
        goto done;
    }
    double* f = last ? fb : fa;
    orc_moments(lat, f, n, F, rho, vel, 1);
    for (size_t i = 0; i < vcells; ++i)
        if (!isfinite(rho[i])) {
            rc = 5;
            goto done;
        }
done:
    free(fa);
    free(fb);
    free(rho);
    free(vel);
    return rc;
}

int main(void)
{
    /* face types: 0 MovingWall, 1 NoSlip, 2 Inflow, 3 Outflow, 4 PressureOutlet, 5 Periodic */
    const int sets[5][6] = {{5, 5, 5, 5, 5, 5}, {1, 1, 1, 0, 1, 1}, {2, 4, 1, 1, 1, 1},
                            {2, 3, 5, 5, 5, 5}, {5, 5, 1, 0, 1, 1}};
    double fu[18] = {0};
    fu[0] = 0.03;          /* inflow velocity on xlo */
    fu[9] = 0.05;          /* lid velocity on yhi */
    const double F[3] = {1e-5, -2e-6, 3e-6};
    const int n2[3] = {9, 6, 1}, n3[3] = {7, 5, 4};
    for (int lat = 0; lat < 4; ++lat)
        for (int s = 0; s < 5; ++s)
            for (int forced = 0; forced < 2; ++forced) {
                const int rc = run_case(lat, lat == 0 ? n2 : n3, sets[s], fu, forced ? F : NULL);
                if (rc) {
                    fprintf(stderr, "lattice %d set %d forced %d: failed (%d)\n", lat, s, forced, rc);
                    return rc;
                }
            }
    printf("oracle sanitize driver: ok\n");
    return 0;
}
