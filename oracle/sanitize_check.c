/* Sanitizer driver of the CPU oracle (test infrastructure only).
 *
 * Built with -fsanitize=address,undefined by `make -C oracle sanitize`
 * (oracle/_asan/orc_check) and run by tests/test_sanitize_cpu.py: it calls
 * every entry point of lgf_oracle.c on small lattices with extents 2 to 6
 * (where neighbour wraps and the 1x3 clover reach cross the whole lattice),
 * so that an out-of-bounds access or undefined arithmetic in the restatement
 * stops the program with a report.  It checks nothing numerical itself
 * beyond finiteness; the numbers are pinned by tests/test_oracle.py. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

void orc_exp_ah(const double *X, double *out);
void orc_staple(const int *shape, const double *U, long g, int mu, double *A);
void orc_flow_gradient(const int *shape, const double *U, double *Z);
void orc_rk3_step(const int *shape, double *U, double eps);
void orc_field_strength(const int *shape, const double *U, long g, int mu, int nu, int kind, double *F);
double orc_energy_density_kind(const int *shape, const double *U, int kind);
double orc_top_charge_kind(const int *shape, const double *U, int kind, double *qx, double *abs_sum);
double orc_avg_plaquette(const int *shape, const double *U);
void orc_unitarity(const int *shape, const double *U, double *maxdev, double *meandev);
int orc_flow_run(const int *shape, double *U, double eps, int nsteps, int every, double *t_out, double *E_out,
                 double *Q_out);
void orc_generate(const int *shape, double *U, double beta, uint64_t seed, int start, int first_sweep, int sweeps,
                  int or_per_hb);
void orc_reunitarize(double *m);
void orc_set_threads(int n);

static int check(const char *what, double v) {
    if (!isfinite(v)) {
        fprintf(stderr, "orc_check: %s is not finite\n", what);
        return 1;
    }
    return 0;
}

int main(void) {
    static const int shapes[][4] = {{4, 4, 4, 4}, {2, 4, 6, 2}, {6, 2, 2, 4}};
    int bad = 0;
    orc_set_threads(2);
    for (int k = 0; k < 3; ++k) {
        const int *sh = shapes[k];
        const long vol = (long)sh[0] * sh[1] * sh[2] * sh[3];
        double *U = malloc(sizeof(double) * vol * 72), *Z = malloc(sizeof(double) * vol * 72);
        double *qx = malloc(sizeof(double) * vol);
        orc_generate(sh, U, 6.0, 1234 + k, 1, 0, 2, 2);  /* hot start, 2 heatbath/OR iterations */
        orc_generate(sh, U, 6.0, 1234 + k, -1, 2, 1, 1); /* continued chain */
        double A[18], F[18], E[18];
        for (long g = 0; g < vol; g += vol / 5 + 1)
            for (int mu = 0; mu < 4; ++mu) {
                orc_staple(sh, U, g, mu, A);
                for (int nu = 0; nu < 4; ++nu)
                    for (int kind = 0; kind < 5 && nu != mu; ++kind) {
                        orc_field_strength(sh, U, g, mu, nu, kind, F);
                        bad |= check("field strength", F[1]);
                    }
            }
        orc_flow_gradient(sh, U, Z);
        for (int i = 0; i < 18; ++i) A[i] = 0.01 * Z[i];
        orc_exp_ah(A, E);
        for (int i = 0; i < 18; ++i) A[i] = 0.4 * Z[18 + i]; /* a large exponent: the other branches */
        orc_exp_ah(A, E);
        bad |= check("exp", E[0]);
        orc_reunitarize(E);
        double t[6], Es[6], Qs[6], abs_sum, mx, mn;
        int n = orc_flow_run(sh, U, 0.02, 4, 2, t, Es, Qs);
        for (int i = 0; i < n; ++i) bad |= check("E", Es[i]) | check("Q", Qs[i]);
        orc_rk3_step(sh, U, 0.02);
        for (int kind = 0; kind <= 4; kind += 4) { /* 1x1 and improved field strength */
            bad |= check("energy", orc_energy_density_kind(sh, U, kind));
            bad |= check("charge", orc_top_charge_kind(sh, U, kind, qx, &abs_sum));
        }
        bad |= check("plaquette", orc_avg_plaquette(sh, U));
        orc_unitarity(sh, U, &mx, &mn);
        bad |= check("unitarity", mx);
        free(U);
        free(Z);
        free(qx);
    }
    printf("orc_check: %s\n", bad ? "FAILED" : "ok");
    return bad;
}
