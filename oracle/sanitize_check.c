/* Test infrastructure: runs every function of the CPU oracle on tiny ragged
 * problems, built with AddressSanitizer and UndefinedBehaviorSanitizer
 * (make -C oracle sanitize).  Any out-of-bounds access, signed overflow or
 * other undefined behaviour aborts the run; the results are only checked
 * for being finite (their values are pinned by the oracle's own tests).
 * The oracle is compiled into this file as one translation unit. */
#include "mt_oracle.c"

#include <stdio.h>

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static double urand(void)
{
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (double)(rng >> 11) * (1.0 / 9007199254740992.0);
}

static int finite_f(const float *a, int64_t n)
{
    for (int64_t i = 0; i < n; i++)
        if (!isfinite(a[i])) return 0;
    return 1;
}
static int finite_d(const double *a, int64_t n)
{
    for (int64_t i = 0; i < n; i++)
        if (!isfinite(a[i])) return 0;
    return 1;
}

/* Tilt geometry with an in-plane axis at phi, centred on the box, plus a
 * detector narrower than the box so that rays and voxels fall off its
 * edges; general = 1 perturbs every coefficient (arbitrary affine views). */
static void geometry(const or_vol *v, const or_prj *p, double phi, int general, double *xc, double *yc)
{
    const double cx = 0.5 * (v->Xstart + v->Xend - 1), cy = 0.5 * (v->Ystart + v->Yend - 1),
                 cz = 0.5 * (v->Zstart + v->Zend - 1);
    for (int a = 0; a < p->AngN; a++) {
        const double th = (-60.0 + 120.0 * a / (p->AngN > 1 ? p->AngN - 1 : 1)) * M_PI / 180.0;
        const double ux = cos(th) * cos(phi), uy = cos(th) * sin(phi), uz = sin(th);
        const double vx = -sin(phi), vy = cos(phi), vz = 0.0;
        double *x = xc + 4 * a, *y = yc + 4 * a;
        x[1] = ux; x[2] = uy; x[3] = uz;
        y[1] = vx; y[2] = vy; y[3] = vz;
        if (general)
            for (int m = 1; m < 4; m++) {
                x[m] += 0.2 * (urand() - 0.5);
                y[m] += 0.2 * (urand() - 0.5);
            }
        x[0] = 0.5 * (p->X + 1) - (x[1] * cx + x[2] * cy + x[3] * cz);
        y[0] = 0.5 * (p->Y + 1) - (y[1] * cx + y[2] * cy + y[3] * cz);
    }
}

static int run(const or_vol *v, const or_prj *p, double phi, int general)
{
    const int64_t N = (int64_t)v->X * v->Y * v->Z, P = (int64_t)p->X * p->Y * p->AngN;
    double *xc = calloc(4 * p->AngN, sizeof(double)), *yc = calloc(4 * p->AngN, sizeof(double));
    float *x = malloc(sizeof(float) * N), *x2 = malloc(sizeof(float) * N), *pr = malloc(sizeof(float) * P);
    float *s = malloc(sizeof(float) * P), *c = malloc(sizeof(float) * P), *r = malloc(sizeof(float) * P);
    double *sd = malloc(sizeof(double) * P), *cd = malloc(sizeof(double) * P), *rd = malloc(sizeof(double) * P);
    double *xd = malloc(sizeof(double) * N), *Sd = malloc(sizeof(double) * N), *Cd = malloc(sizeof(double) * N);
    float *z = malloc(sizeof(float) * N);
    int ok = xc && yc && x && x2 && pr && s && c && r && sd && cd && rd && xd && Sd && Cd && z;
    if (ok) {
        geometry(v, p, phi, general, xc, yc);
        for (int64_t i = 0; i < N; i++) x[i] = (float)urand();
        for (int64_t i = 0; i < P; i++) pr[i] = (float)(10.0 * urand());
        int i0, k0;
        double du, dv;
        or_weight(v->Xstart, v->Ystart, v->Zend - 1, p->AngN - 1, xc, yc, &i0, &k0, &du, &dv);
        ok = ok && isfinite(du) && isfinite(dv);
        ok = ok && or_fp(p, v, xc, yc, x, s, c) == 0 && finite_f(s, P) && finite_f(c, P);
        ok = ok && or_residual(P, pr, s, c, r) == 0 && finite_f(r, P);
        const int zm = v->Zstart + v->Z / 2;
        ok = ok && or_fp_slab(p, v, v->Zstart, zm, xc, yc, x, s, c) == 0;
        memcpy(x2, x, sizeof(float) * N);
        ok = ok && or_bp(p, v, xc, yc, r, x2, 0.5f) == 0 && finite_f(x2, N);
        ok = ok && or_bp_slab(p, v, zm, v->Zend, xc, yc, r, x2, 0.5f) == 0 && finite_f(x2, N);
        ok = ok && or_bpt(p, v, xc, yc, pr, x2) == 0 && finite_f(x2, N);
        ok = ok && or_sirt(p, v, xc, yc, pr, x2, 2, 0.7f) == 0 && finite_f(x2, N);
        for (int64_t i = 0; i < N; i++) xd[i] = x[i];
        ok = ok && or_fp_f64(p, v, xc, yc, xd, sd, cd) == 0 && finite_d(sd, P) && finite_d(cd, P);
        ok = ok && or_residual_f64(P, pr, sd, cd, rd) == 0 && finite_d(rd, P);
        ok = ok && or_bp_sums_volume_f64(p, v, xc, yc, rd, Sd, Cd) == 0 && finite_d(Sd, N) && finite_d(Cd, N);
        ok = ok && or_bp_f64(p, v, xc, yc, rd, xd, 0.5) == 0 && finite_d(xd, N);
        ok = ok && or_bpt_f64(p, v, xc, yc, pr, xd) == 0 && finite_d(xd, N);
        ok = ok && or_sirt_f64(p, v, xc, yc, pr, xd, 2, 0.7) == 0 && finite_d(xd, N);
        /* MACE over two agents sharing the box */
        const or_prj pp[2] = {*p, *p};
        const double *xcs[2] = {xc, xc}, *ycs[2] = {yc, yc};
        const float *ps[2] = {pr, pr};
        ok = ok && or_mace(2, pp, v, xcs, ycs, ps, 2, 2, 0.5f, 0.5f, x2) == 0 && finite_f(x2, N);
        ok = ok && or_mace_f64(2, pp, v, xcs, ycs, ps, 2, 2, 0.5, 0.5, xd) == 0 && finite_d(xd, N);
        float *ws[2] = {x, x2};
        ok = ok && or_form_z(N, x, x2, z) == 0 && or_mann(N, x, x2, z, 0.5f) == 0;
        ok = ok && or_mean(N, 2, (const float *const *)ws, x) == 0 && finite_f(x, N);
        for (int id = 0; id < 5; id++) {
            const or_slice sl = or_divide_slice(id, 5, v->Z, v->Zstart);
            ok = ok && sl.z_start <= sl.z_end;
        }
    }
    free(xc); free(yc); free(x); free(x2); free(pr); free(s); free(c); free(r);
    free(sd); free(cd); free(rd); free(xd); free(Sd); free(Cd); free(z);
    return ok;
}

int main(void)
{
    const or_vol vols[2] = {{-3, 4, 2, 7, -1, 3, 7, 5, 4}, {0, 6, 0, 9, 0, 5, 6, 9, 5}};
    const or_prj prjs[2] = {{5, 4, 5}, {11, 10, 3}};
    const double phis[3] = {0.0, 0.5 * M_PI, 0.25 * M_PI};
    int fails = 0, runs = 0;
    for (int g = 0; g < 2; g++)
        for (int q = 0; q < 2; q++)
            for (int f = 0; f < 3; f++) {
                runs++;
                if (!run(&vols[q], &prjs[q], phis[f], g)) {
                    fprintf(stderr, "case vol %d prj %d phi %d general %d failed\n", q, q, f, g);
                    fails++;
                }
            }
    printf("%d of %d cases clean\n", runs - fails, runs);
    return fails ? 1 : 0;
}
