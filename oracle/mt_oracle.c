/*
 * mt_oracle.c -- CPU restatement of the reference's SIRT + MACE arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * product path and the timed CPU baseline.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it; the
 * product library never links, calls or falls back to it.
 *
 * Two arithmetic modes are restated:
 *   "ref"  -- the reference's own mixed precision: FP64 coordinates and
 *             back-projection sums, FP32 splat contributions, FP32 residual,
 *             Mann update, z-formation and consensus mean.
 *   "f64"  -- the same operator with every value held in double; the FP64
 *             oracle that split-precision results are measured against.
 *
 * Reference citations are file:line in Autumn1998/Multi-tilt-Reconstuction.
 * Parity pinning of each function (see tests/test_oracle_*.py):
 *   or_weight        closed form on integer-aligned geometry; long-double
 *                    tent-function brute force on tiny random geometry.
 *   or_fp / or_bp    dense tent-function matrix brute force (independent
 *                    formulation), adjointness, constant-field and
 *                    integer-shift known answers.
 *   or_residual      known answer (c == 0 branch and s/c branch).
 *   or_sirt          fixed point, affinity in x0 (f64 mode), trivial
 *                    one-to-one geometry closed form.
 *   or_mace          K=1/rho=0.5 reduction to SIRT^(S*M)(BPT), rho=0,
 *                    K identical agents, hand-computed K=2 closed form.
 *   or_divide_slice  partition invariants (contiguous, covering, sizes).
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Layout-identical to the reference's VolumeSize / ProjectionSize
 * (MrcFileIOBaseFunction.h:57-70, :73-78). */
typedef struct {
    int Xstart, Xend, Ystart, Yend, Zstart, Zend;
    int X, Y, Z;
} or_vol;

typedef struct {
    int X, Y, AngN;
} or_prj;

/* Slice (MrcFileIOBaseFunction.h:50-54). */
typedef struct {
    int z_start, z_end;
} or_slice;

/* ------------------------------------------------------------------ */
/* a1: computeWeight, FunctionOnGPU.cu:12-25.
 * u = xc0 + xc1*X + xc2*Y + xc3*Z - 1 evaluated left to right in double;
 * CUDA contracts the chain into fused multiply-adds, restated here with
 * explicit fma() so the oracle does not depend on the host compiler's
 * contraction choices. */
static inline void or_weight_impl(int X, int Y, int Z, int a,
                                  const double *xc, const double *yc,
                                  int *i0, int *k0, double *du, double *dv)
{
    const int q = 4 * a;
    double u = fma(xc[q + 3], (double)Z, fma(xc[q + 2], (double)Y, fma(xc[q + 1], (double)X, xc[q]))) - 1.0;
    double v = fma(yc[q + 3], (double)Z, fma(yc[q + 2], (double)Y, fma(yc[q + 1], (double)X, yc[q]))) - 1.0;
    int iu = (int)floor(u);
    int iv = (int)floor(v);
    *i0 = iu;
    *k0 = iv;
    *du = u - (double)iu;
    *dv = v - (double)iv;
}

void or_weight(int X, int Y, int Z, int a, const double *xc, const double *yc,
               int *i0, int *k0, double *du, double *dv)
{
    or_weight_impl(X, Y, Z, a, xc, yc, i0, k0, du, dv);
}

/* ------------------------------------------------------------------ */
/* a15: divideSlice, MultiAxis.cpp:13-31.  The first Z%N ranks get one
 * extra slice. */
or_slice or_divide_slice(int myid, int N, int Z, int S)
{
    or_slice sl;
    int zrem = Z % N;
    int height;
    if (myid < zrem) {
        height = Z / N + 1;
        sl.z_start = height * myid + S;
    } else {
        height = Z / N;
        sl.z_start = height * myid + zrem + S;
    }
    sl.z_end = sl.z_start + height;
    return sl;
}

/* ------------------------------------------------------------------ */
/* a2 + a4: computeDivisor (FunctionOnGPU.cu:316-333) calling Reproj
 * (FunctionOnGPU.cu:27-52).  Voxel-driven bilinear splat: each in-range
 * tap adds fl32(w * x) to s and fl32(w) to c.  The reference sums with
 * float atomics in an undefined order; this restatement sums each view's
 * pixels in voxel memory order (z, y, x), one view per thread, so the
 * result is deterministic for any thread count.  s and c are zeroed here
 * (the reference zeroes them before each launch, FunctionOnGPU.cu:214-217).
 * Slab [zs, ze) of the volume is projected (the reference's per-rank slab,
 * MultiAxis.cpp:163); x holds that slab, [z-zs][y][x]. */
int or_fp_slab(const or_prj *prj, const or_vol *vol, int zs, int ze,
               const double *xc, const double *yc, const float *x,
               float *s, float *c)
{
    const int nu = prj->X, nv = prj->Y, A = prj->AngN;
    const int64_t plane = (int64_t)nu * nv;
    memset(s, 0, sizeof(float) * plane * A);
    memset(c, 0, sizeof(float) * plane * A);
#pragma omp parallel for schedule(dynamic, 1)
    for (int a = 0; a < A; a++) {
        float *sv = s + plane * a;
        float *cv = c + plane * a;
        for (int Z = zs; Z < ze; Z++)
            for (int Y = vol->Ystart; Y < vol->Yend; Y++)
                for (int X = vol->Xstart; X < vol->Xend; X++) {
                    const int64_t j = (int64_t)(X - vol->Xstart) + (int64_t)(Y - vol->Ystart) * vol->X
                                    + (int64_t)(Z - zs) * vol->X * vol->Y;
                    const double xv = (double)x[j];
                    int i0, k0;
                    double du, dv;
                    or_weight_impl(X, Y, Z, a, xc, yc, &i0, &k0, &du, &dv);
                    const int in0 = (i0 >= 0 && i0 < nu), in1 = (i0 + 1 >= 0 && i0 + 1 < nu);
                    const int kn0 = (k0 >= 0 && k0 < nv), kn1 = (k0 + 1 >= 0 && k0 + 1 < nv);
                    if (in0 && kn0) {
                        double w = (1 - du) * (1 - dv);
                        int64_t n = (int64_t)i0 + (int64_t)k0 * nu;
                        sv[n] += (float)(w * xv);
                        cv[n] += (float)w;
                    }
                    if (in1 && kn0) {
                        double w = du * (1 - dv);
                        int64_t n = (int64_t)(i0 + 1) + (int64_t)k0 * nu;
                        sv[n] += (float)(w * xv);
                        cv[n] += (float)w;
                    }
                    if (in0 && kn1) {
                        double w = (1 - du) * dv;
                        int64_t n = (int64_t)i0 + (int64_t)(k0 + 1) * nu;
                        sv[n] += (float)(w * xv);
                        cv[n] += (float)w;
                    }
                    if (in1 && kn1) {
                        double w = du * dv;
                        int64_t n = (int64_t)(i0 + 1) + (int64_t)(k0 + 1) * nu;
                        sv[n] += (float)(w * xv);
                        cv[n] += (float)w;
                    }
                }
    }
    return 0;
}

int or_fp(const or_prj *prj, const or_vol *vol, const double *xc, const double *yc,
          const float *x, float *s, float *c)
{
    return or_fp_slab(prj, vol, vol->Zstart, vol->Zend, xc, yc, x, s, c);
}

/* a5: residual on the group leader, FunctionOnGPU.cu:228-236 (FP32):
 * r = p - (c != 0 ? s / c : s). */
int or_residual(int64_t P, const float *p, const float *s, const float *c, float *r)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < P; k++) {
        float sk = s[k];
        if (c[k] != 0) sk /= c[k];
        r[k] = p[k] - sk;
    }
    return 0;
}

/* a3 + a6 / a7: BilinearValue (FunctionOnGPU.cu:54-79) accumulated over
 * views in order 0..A-1 in double (FunctionOnGPU.cu:113-130).  Returns S
 * and C for voxel (X, Y, Z). */
static inline void or_bp_sums(const or_prj *prj, int X, int Y, int Z,
                              const double *xc, const double *yc, const float *r,
                              double *S, double *C)
{
    const int nu = prj->X, nv = prj->Y, A = prj->AngN;
    const int64_t plane = (int64_t)nu * nv;
    double s = 0, c = 0;
    for (int a = 0; a < A; a++) {
        int i0, k0;
        double du, dv;
        or_weight_impl(X, Y, Z, a, xc, yc, &i0, &k0, &du, &dv);
        const float *rv = r + plane * a;
        const int in0 = (i0 >= 0 && i0 < nu), in1 = (i0 + 1 >= 0 && i0 + 1 < nu);
        const int kn0 = (k0 >= 0 && k0 < nv), kn1 = (k0 + 1 >= 0 && k0 + 1 < nv);
        if (in0 && kn0) {
            double w = (1 - du) * (1 - dv);
            s = fma(w, (double)rv[i0 + (int64_t)k0 * nu], s);
            c += w;
        }
        if (in1 && kn0) {
            double w = du * (1 - dv);
            s = fma(w, (double)rv[i0 + 1 + (int64_t)k0 * nu], s);
            c += w;
        }
        if (in0 && kn1) {
            double w = (1 - du) * dv;
            s = fma(w, (double)rv[i0 + (int64_t)(k0 + 1) * nu], s);
            c += w;
        }
        if (in1 && kn1) {
            double w = du * dv;
            s = fma(w, (double)rv[i0 + 1 + (int64_t)(k0 + 1) * nu], s);
            c += w;
        }
    }
    *S = s;
    *C = c;
}

/* a6: sirtBackProjOnGPU, FunctionOnGPU.cu:111-135:
 * if C != 0: x = fl32((double)x + (double)fl32(S/C) * lambda).
 * lambda is the float Options::sirt_step (ReadAndInitialOptions.h:29)
 * promoted to the kernel's double argument. */
int or_bp_slab(const or_prj *prj, const or_vol *vol, int zs, int ze,
               const double *xc, const double *yc, const float *r, float *x, float lambda)
{
    const double lam = (double)lambda;
    const int nz = ze - zs;
#pragma omp parallel for collapse(2) schedule(static)
    for (int zz = 0; zz < nz; zz++)
        for (int Y = vol->Ystart; Y < vol->Yend; Y++) {
            const int Z = zs + zz;
            for (int X = vol->Xstart; X < vol->Xend; X++) {
                const int64_t j = (int64_t)(X - vol->Xstart) + (int64_t)(Y - vol->Ystart) * vol->X
                                + (int64_t)zz * vol->X * vol->Y;
                double S, C;
                or_bp_sums(prj, X, Y, Z, xc, yc, r, &S, &C);
                if (C != 0.0) x[j] = (float)fma((double)(float)(S / C), lam, (double)x[j]);
            }
        }
    return 0;
}

int or_bp(const or_prj *prj, const or_vol *vol, const double *xc, const double *yc,
          const float *r, float *x, float lambda)
{
    return or_bp_slab(prj, vol, vol->Zstart, vol->Zend, xc, yc, r, x, lambda);
}

/* a7 + a9: backProjOnGPU (FunctionOnGPU.cu:82-109) as driven by
 * ComputeInitialModel (FunctionOnGPU.cu:138-179): x += (float)(S/C) with
 * r := p.  The reference leaves the device buffer unzeroed
 * (FunctionOnGPU.cu:157); the restatement zeroes it. */
int or_bpt(const or_prj *prj, const or_vol *vol, const double *xc, const double *yc,
           const float *p, float *x)
{
    const int64_t N = (int64_t)vol->X * vol->Y * vol->Z;
#pragma omp parallel for collapse(2) schedule(static)
    for (int zz = 0; zz < vol->Z; zz++)
        for (int Y = vol->Ystart; Y < vol->Yend; Y++) {
            const int Z = vol->Zstart + zz;
            for (int X = vol->Xstart; X < vol->Xend; X++) {
                const int64_t j = (int64_t)(X - vol->Xstart) + (int64_t)(Y - vol->Ystart) * vol->X
                                + (int64_t)zz * vol->X * vol->Y;
                double S, C;
                or_bp_sums(prj, X, Y, Z, xc, yc, p, &S, &C);
                float acc = 0.0f;
                if (C != 0.0) acc += (float)(S / C);
                x[j] = acc;
            }
        }
    (void)N;
    return 0;
}

/* a8: runSIRTOnGPU loop, FunctionOnGPU.cu:212-244: iters x {FP; residual;
 * BP}.  The reference hard-codes iters = 10 (FunctionOnGPU.cu:212). */
int or_sirt(const or_prj *prj, const or_vol *vol, const double *xc, const double *yc,
            const float *p, float *x, int iters, float lambda)
{
    const int64_t P = (int64_t)prj->X * prj->Y * prj->AngN;
    float *s = (float *)malloc(sizeof(float) * P);
    float *c = (float *)malloc(sizeof(float) * P);
    float *r = (float *)malloc(sizeof(float) * P);
    if (!s || !c || !r) { free(s); free(c); free(r); return -1; }
    for (int it = 0; it < iters; it++) {
        or_fp(prj, vol, xc, yc, x, s, c);
        or_residual(P, p, s, c, r);
        or_bp(prj, vol, xc, yc, r, x, lambda);
    }
    free(s); free(c); free(r);
    return 0;
}

/* a10: initialZOnGPU, FunctionOnGPU.cu:268-276: z = 2*wbar - w (FP32). */
int or_form_z(int64_t N, const float *w, const float *wbar, float *z)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) z[i] = 2 * wbar[i] - w[i];
    return 0;
}

/* a11: runMaceOnCPU, MultiAxis.cpp:33-38: w = rho*(2v - z) + (1-rho)*w
 * (FP32, host code, not contracted). */
int or_mann(int64_t N, const float *v, float *w, const float *z, float rho)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) w[i] = rho * (2 * v[i] - z[i]) + (1 - rho) * w[i];
    return 0;
}

/* a12: consensus mean, MultiAxis.cpp:61-66: wbar = (sum over agents of
 * w_a) / K, summed in agent order (the MPI_SUM order is implementation
 * defined; with one rank per group numprocs == K). */
int or_mean(int64_t N, int K, const float *const *w, float *wbar)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; i++) {
        float acc = w[0][i];
        for (int a = 1; a < K; a++) acc += w[a][i];
        wbar[i] = acc / (float)K;
    }
    return 0;
}

/* a13 + a14: BPT-sum initialisation (MultiAxis.cpp:163-200) followed by
 * reconstructModel (MultiAxis.cpp:40-71).  All agents share the volume
 * vol (synchronizeVol, MultiAxis.cpp:73-99).  Output: wbar.
 *   B = sum_a BPT_a;  w_a = B;  wbar = B   (wbar aliases initial_model,
 *   MultiAxis.cpp:200,203)
 *   M times: z_a = 2 wbar - w_a; v_a = SIRT^S(z_a); w_a = Mann(v_a, z_a, w_a);
 *            wbar = mean_a w_a. */
int or_mace(int K, const or_prj *prj, const or_vol *vol,
            const double *const *xc, const double *const *yc, const float *const *p,
            int M, int S, float lambda, float rho, float *wbar)
{
    const int64_t N = (int64_t)vol->X * vol->Y * vol->Z;
    float **w = (float **)calloc(K, sizeof(float *));
    float *z = (float *)malloc(sizeof(float) * N);
    float *tmp = (float *)malloc(sizeof(float) * N);
    int rc = 0;
    if (!w || !z || !tmp) { rc = -1; goto done; }
    for (int a = 0; a < K; a++) {
        w[a] = (float *)malloc(sizeof(float) * N);
        if (!w[a]) { rc = -1; goto done; }
    }
    /* B = sum over agents of BPT (MPI_Reduce_BIG SUM, MultiAxis.cpp:192) */
    for (int64_t i = 0; i < N; i++) wbar[i] = 0.0f;
    for (int a = 0; a < K; a++) {
        or_bpt(&prj[a], vol, xc[a], yc[a], p[a], tmp);
        for (int64_t i = 0; i < N; i++) wbar[i] += tmp[i];
    }
    for (int a = 0; a < K; a++) memcpy(w[a], wbar, sizeof(float) * N);
    for (int m = 0; m < M; m++) {
        for (int a = 0; a < K; a++) {
            or_form_z(N, w[a], wbar, z);
            memcpy(tmp, z, sizeof(float) * N);
            or_sirt(&prj[a], vol, xc[a], yc[a], p[a], tmp, S, lambda);
            or_mann(N, tmp, w[a], z, rho);
        }
        or_mean(N, K, (const float *const *)w, wbar);
    }
done:
    if (w) for (int a = 0; a < K; a++) free(w[a]);
    free(w); free(z); free(tmp);
    return rc;
}

/* ================================================================== */
/* f64 mode: the same operator with every quantity in double.          */
/* ================================================================== */

int or_fp_f64(const or_prj *prj, const or_vol *vol, const double *xc, const double *yc,
              const double *x, double *s, double *c)
{
    const int nu = prj->X, nv = prj->Y, A = prj->AngN;
    const int64_t plane = (int64_t)nu * nv;
    memset(s, 0, sizeof(double) * plane * A);
    memset(c, 0, sizeof(double) * plane * A);
#pragma omp parallel for schedule(dynamic, 1)
    for (int a = 0; a < A; a++) {
        double *sv = s + plane * a;
        double *cv = c + plane * a;
        for (int Z = vol->Zstart; Z < vol->Zend; Z++)
            for (int Y = vol->Ystart; Y < vol->Yend; Y++)
                for (int X = vol->Xstart; X < vol->Xend; X++) {
                    const int64_t j = (int64_t)(X - vol->Xstart) + (int64_t)(Y - vol->Ystart) * vol->X
                                    + (int64_t)(Z - vol->Zstart) * vol->X * vol->Y;
                    const double xv = x[j];
                    int i0, k0;
                    double du, dv;
                    or_weight_impl(X, Y, Z, a, xc, yc, &i0, &k0, &du, &dv);
                    const int in0 = (i0 >= 0 && i0 < nu), in1 = (i0 + 1 >= 0 && i0 + 1 < nu);
                    const int kn0 = (k0 >= 0 && k0 < nv), kn1 = (k0 + 1 >= 0 && k0 + 1 < nv);
                    if (in0 && kn0) { double w = (1 - du) * (1 - dv); int64_t n = i0 + (int64_t)k0 * nu; sv[n] += w * xv; cv[n] += w; }
                    if (in1 && kn0) { double w = du * (1 - dv); int64_t n = i0 + 1 + (int64_t)k0 * nu; sv[n] += w * xv; cv[n] += w; }
                    if (in0 && kn1) { double w = (1 - du) * dv; int64_t n = i0 + (int64_t)(k0 + 1) * nu; sv[n] += w * xv; cv[n] += w; }
                    if (in1 && kn1) { double w = du * dv; int64_t n = i0 + 1 + (int64_t)(k0 + 1) * nu; sv[n] += w * xv; cv[n] += w; }
                }
    }
    return 0;
}

int or_residual_f64(int64_t P, const float *p, const double *s, const double *c, double *r)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < P; k++) {
        double sk = s[k];
        if (c[k] != 0) sk /= c[k];
        r[k] = (double)p[k] - sk;
    }
    return 0;
}

static inline void or_bp_sums_f64(const or_prj *prj, int X, int Y, int Z,
                                  const double *xc, const double *yc, const double *r,
                                  double *S, double *C)
{
    const int nu = prj->X, nv = prj->Y, A = prj->AngN;
    const int64_t plane = (int64_t)nu * nv;
    double s = 0, c = 0;
    for (int a = 0; a < A; a++) {
        int i0, k0;
        double du, dv;
        or_weight_impl(X, Y, Z, a, xc, yc, &i0, &k0, &du, &dv);
        const double *rv = r + plane * a;
        const int in0 = (i0 >= 0 && i0 < nu), in1 = (i0 + 1 >= 0 && i0 + 1 < nu);
        const int kn0 = (k0 >= 0 && k0 < nv), kn1 = (k0 + 1 >= 0 && k0 + 1 < nv);
        if (in0 && kn0) { double w = (1 - du) * (1 - dv); s += w * rv[i0 + (int64_t)k0 * nu]; c += w; }
        if (in1 && kn0) { double w = du * (1 - dv); s += w * rv[i0 + 1 + (int64_t)k0 * nu]; c += w; }
        if (in0 && kn1) { double w = (1 - du) * dv; s += w * rv[i0 + (int64_t)(k0 + 1) * nu]; c += w; }
        if (in1 && kn1) { double w = du * dv; s += w * rv[i0 + 1 + (int64_t)(k0 + 1) * nu]; c += w; }
    }
    *S = s;
    *C = c;
}

/* Un-normalised back-projection S = A^T r and C = A^T 1 (f64); used by the
 * adjointness checks. */
int or_bp_sums_volume_f64(const or_prj *prj, const or_vol *vol, const double *xc, const double *yc,
                          const double *r, double *S, double *C)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int zz = 0; zz < vol->Z; zz++)
        for (int Y = vol->Ystart; Y < vol->Yend; Y++)
            for (int X = vol->Xstart; X < vol->Xend; X++) {
                const int64_t j = (int64_t)(X - vol->Xstart) + (int64_t)(Y - vol->Ystart) * vol->X
                                + (int64_t)zz * vol->X * vol->Y;
                or_bp_sums_f64(prj, X, Y, vol->Zstart + zz, xc, yc, r, &S[j], &C[j]);
            }
    return 0;
}

int or_bp_f64(const or_prj *prj, const or_vol *vol, const double *xc, const double *yc,
              const double *r, double *x, double lambda)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int zz = 0; zz < vol->Z; zz++)
        for (int Y = vol->Ystart; Y < vol->Yend; Y++)
            for (int X = vol->Xstart; X < vol->Xend; X++) {
                const int64_t j = (int64_t)(X - vol->Xstart) + (int64_t)(Y - vol->Ystart) * vol->X
                                + (int64_t)zz * vol->X * vol->Y;
                double S, C;
                or_bp_sums_f64(prj, X, Y, vol->Zstart + zz, xc, yc, r, &S, &C);
                if (C != 0.0) x[j] += (S / C) * lambda;
            }
    return 0;
}

int or_bpt_f64(const or_prj *prj, const or_vol *vol, const double *xc, const double *yc,
               const float *p, double *x)
{
    const int64_t P = (int64_t)prj->X * prj->Y * prj->AngN;
    const int64_t N = (int64_t)vol->X * vol->Y * vol->Z;
    double *pd = (double *)malloc(sizeof(double) * P);
    if (!pd) return -1;
    for (int64_t k = 0; k < P; k++) pd[k] = (double)p[k];
    for (int64_t i = 0; i < N; i++) x[i] = 0.0;
    or_bp_f64(prj, vol, xc, yc, pd, x, 1.0);
    free(pd);
    return 0;
}

int or_sirt_f64(const or_prj *prj, const or_vol *vol, const double *xc, const double *yc,
                const float *p, double *x, int iters, double lambda)
{
    const int64_t P = (int64_t)prj->X * prj->Y * prj->AngN;
    double *s = (double *)malloc(sizeof(double) * P);
    double *c = (double *)malloc(sizeof(double) * P);
    double *r = (double *)malloc(sizeof(double) * P);
    if (!s || !c || !r) { free(s); free(c); free(r); return -1; }
    for (int it = 0; it < iters; it++) {
        or_fp_f64(prj, vol, xc, yc, x, s, c);
        or_residual_f64(P, p, s, c, r);
        or_bp_f64(prj, vol, xc, yc, r, x, lambda);
    }
    free(s); free(c); free(r);
    return 0;
}

int or_mace_f64(int K, const or_prj *prj, const or_vol *vol,
                const double *const *xc, const double *const *yc, const float *const *p,
                int M, int S, double lambda, double rho, double *wbar)
{
    const int64_t N = (int64_t)vol->X * vol->Y * vol->Z;
    double **w = (double **)calloc(K, sizeof(double *));
    double *z = (double *)malloc(sizeof(double) * N);
    double *tmp = (double *)malloc(sizeof(double) * N);
    int rc = 0;
    if (!w || !z || !tmp) { rc = -1; goto done; }
    for (int a = 0; a < K; a++) {
        w[a] = (double *)malloc(sizeof(double) * N);
        if (!w[a]) { rc = -1; goto done; }
    }
    for (int64_t i = 0; i < N; i++) wbar[i] = 0.0;
    for (int a = 0; a < K; a++) {
        or_bpt_f64(&prj[a], vol, xc[a], yc[a], p[a], tmp);
        for (int64_t i = 0; i < N; i++) wbar[i] += tmp[i];
    }
    for (int a = 0; a < K; a++) memcpy(w[a], wbar, sizeof(double) * N);
    for (int m = 0; m < M; m++) {
        for (int a = 0; a < K; a++) {
            for (int64_t i = 0; i < N; i++) z[i] = 2 * wbar[i] - w[a][i];
            memcpy(tmp, z, sizeof(double) * N);
            or_sirt_f64(&prj[a], vol, xc[a], yc[a], p[a], tmp, S, lambda);
            for (int64_t i = 0; i < N; i++) w[a][i] = rho * (2 * tmp[i] - z[i]) + (1 - rho) * w[a][i];
        }
        for (int64_t i = 0; i < N; i++) {
            double acc = 0;
            for (int a = 0; a < K; a++) acc += w[a][i];
            wbar[i] = acc / K;
        }
    }
done:
    if (w) for (int a = 0; a < K; a++) free(w[a]);
    free(w); free(z); free(tmp);
    return rc;
}

/* Thread-count control for the timed CPU baseline. */
int or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
