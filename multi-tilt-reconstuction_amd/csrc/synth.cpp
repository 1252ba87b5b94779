// Seeded synthetic tilt-series generator: phantom, geometry, analytic ray
// sums and noise.  Host code (std::thread), linked into libmtr.so.
#include "mtr_synth.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

struct Xoshiro256 {
    uint64_t s[4];
    explicit Xoshiro256(uint64_t seed)
    {
        uint64_t sm = seed;
        for (int i = 0; i < 4; i++) s[i] = splitmix64(sm);
    }
    uint64_t next()
    {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // Box-Muller; u1 in (0,1] keeps log finite.
    void normal2(double &a, double &b)
    {
        const double u1 = 1.0 - uniform();
        const double u2 = uniform();
        const double r = std::sqrt(-2.0 * std::log(u1));
        const double t = 6.283185307179586476925286766559 * u2;
        a = r * std::cos(t);
        b = r * std::sin(t);
    }
};

// cos/sin of an angle in degrees, exact at multiples of 90 degrees so that
// axis-aligned geometries carry exact 0 / +-1 coefficients.
void cos_sin_deg(double deg, double &c, double &s)
{
    double m = std::fmod(deg, 360.0);
    if (m < 0) m += 360.0;
    if (m == 0.0) { c = 1; s = 0; return; }
    if (m == 90.0) { c = 0; s = 1; return; }
    if (m == 180.0) { c = -1; s = 0; return; }
    if (m == 270.0) { c = 0; s = -1; return; }
    const double r = deg * 0.017453292519943295769236907684886;
    c = std::cos(r);
    s = std::sin(r);
}

int resolve_threads(int threads)
{
    if (threads > 0) return threads;
    unsigned hc = std::thread::hardware_concurrency();
    return hc ? (int)std::min(hc, 64u) : 1;
}

template <class F>
void parallel_for(int n, int threads, F f)
{
    threads = std::max(1, std::min(threads, n));
    if (threads == 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(threads);
    for (int t = 0; t < threads; t++) {
        auto job = [=]() {
            for (int i = t; i < n; i += threads) f(i);
        };
        try {
            pool.emplace_back(job);
        } catch (...) {
            job();  // no thread to spare: this share here
        }
    }
    for (auto &th : pool) th.join();
}

bool vol_ok(const mtr_volume_size *v)
{
    return v && v->X > 0 && v->Y > 0 && v->Z > 0 && v->Xend - v->Xstart == v->X &&
           v->Yend - v->Ystart == v->Y && v->Zend - v->Zstart == v->Z;
}

}  // namespace

extern "C" int mtr_synth_geometry(const mtr_volume_size *vol, const mtr_projection_size *prj,
                                  int kind, double phi_deg, const double *theta_deg, double *xc,
                                  double *yc)
{
    if (!vol_ok(vol) || !prj || prj->X <= 0 || prj->Y <= 0 || prj->AngN <= 0 || !theta_deg ||
        !xc || !yc)
        return MTR_ERR_ARG;
    const double cx = 0.5 * (vol->Xstart + vol->Xend - 1);
    const double cy = 0.5 * (vol->Ystart + vol->Yend - 1);
    const double cz = 0.5 * (vol->Zstart + vol->Zend - 1);
    const double cu = 0.5 * (prj->X - 1);
    const double cv = 0.5 * (prj->Y - 1);
    double cp, sp;
    cos_sin_deg(phi_deg, cp, sp);
    for (int a = 0; a < prj->AngN; a++) {
        double ct, st;
        cos_sin_deg(theta_deg[a], ct, st);
        double *X = xc + 4 * a, *Y = yc + 4 * a;
        // The stored constant carries +1 so that the reference's "- 1"
        // (FunctionOnGPU.cu:17-18) lands the centre on (cu, cv).
        if (kind == MTR_GEOM_SLAB) {
            X[1] = ct * cp;
            X[2] = ct * sp;
            X[3] = st;
            X[0] = cu + 1.0 - (X[1] * cx + X[2] * cy + X[3] * cz);
            Y[1] = -sp;
            Y[2] = cp;
            Y[3] = 0.0;
            Y[0] = cv + 1.0 - (Y[1] * cx + Y[2] * cy);
        } else if (kind == MTR_GEOM_AXIS_Z) {
            X[1] = ct;
            X[2] = st;
            X[3] = 0.0;
            X[0] = cu + 1.0 - (X[1] * cx + X[2] * cy);
            Y[1] = 0.0;
            Y[2] = 0.0;
            Y[3] = 1.0;
            Y[0] = cv + 1.0 - cz;
        } else {
            return MTR_ERR_ARG;
        }
    }
    return MTR_OK;
}

extern "C" int mtr_synth_phantom(uint64_t seed, const mtr_volume_size *vol, int n, double amin,
                                 double amax, double dmin, double dmax, double *ell)
{
    if (!vol_ok(vol) || n < 0 || !ell || amin <= 0 || amax < amin) return MTR_ERR_ARG;
    Xoshiro256 rng(seed);
    for (int e = 0; e < n; e++) {
        double *E = ell + (size_t)MTR_ELLIPSOID_DOUBLES * e;
        E[0] = vol->Xstart + rng.uniform() * (vol->X - 1);
        E[1] = vol->Ystart + rng.uniform() * (vol->Y - 1);
        E[2] = vol->Zstart + rng.uniform() * (vol->Z - 1);
        for (int k = 0; k < 3; k++) E[3 + k] = amin + rng.uniform() * (amax - amin);
        // Shoemake: uniform random unit quaternion.
        const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
        const double s1 = std::sqrt(1.0 - u1), s2 = std::sqrt(u1);
        const double t2 = 6.283185307179586476925286766559 * u2;
        const double t3 = 6.283185307179586476925286766559 * u3;
        const double qx = s1 * std::sin(t2), qy = s1 * std::cos(t2);
        const double qz = s2 * std::sin(t3), qw = s2 * std::cos(t3);
        double *R = E + 6;
        R[0] = 1 - 2 * (qy * qy + qz * qz);
        R[1] = 2 * (qx * qy - qz * qw);
        R[2] = 2 * (qx * qz + qy * qw);
        R[3] = 2 * (qx * qy + qz * qw);
        R[4] = 1 - 2 * (qx * qx + qz * qz);
        R[5] = 2 * (qy * qz - qx * qw);
        R[6] = 2 * (qx * qz - qy * qw);
        R[7] = 2 * (qy * qz + qx * qw);
        R[8] = 1 - 2 * (qx * qx + qy * qy);
        E[15] = dmin + rng.uniform() * (dmax - dmin);
    }
    return MTR_OK;
}

extern "C" int mtr_synth_project(const mtr_volume_size *vol, const mtr_projection_size *prj,
                                 const double *xc, const double *yc, int n, const double *ell,
                                 float *p, int threads)
{
    if (!vol_ok(vol) || !prj || prj->X <= 0 || prj->Y <= 0 || prj->AngN <= 0 || !xc || !yc ||
        n < 0 || (n > 0 && !ell) || !p)
        return MTR_ERR_ARG;
    const int nu = prj->X, nv = prj->Y;
    const int64_t plane = (int64_t)nu * nv;
    const double C[3] = {0.5 * (vol->Xstart + vol->Xend - 1), 0.5 * (vol->Ystart + vol->Yend - 1),
                         0.5 * (vol->Zstart + vol->Zend - 1)};
    int bad = 0;
    parallel_for(prj->AngN, resolve_threads(threads), [&](int a) {
        float *img = p + plane * a;
        std::vector<double> acc((size_t)plane, 0.0);
        const double *gu = xc + 4 * a + 1, *gv = yc + 4 * a + 1;
        // Ray direction d = grad u x grad v, unit length.
        double d[3] = {gu[1] * gv[2] - gu[2] * gv[1], gu[2] * gv[0] - gu[0] * gv[2],
                       gu[0] * gv[1] - gu[1] * gv[0]};
        const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (!(dn > 0)) { bad = 1; return; }
        for (double &q : d) q /= dn;
        // Minimum-norm point of pixel (i,k): P = C + G^T H^-1 e with
        // e = (i + 1 - xc0 - gu.C, k + 1 - yc0 - gv.C), H = G G^T.
        const double h00 = gu[0] * gu[0] + gu[1] * gu[1] + gu[2] * gu[2];
        const double h01 = gu[0] * gv[0] + gu[1] * gv[1] + gu[2] * gv[2];
        const double h11 = gv[0] * gv[0] + gv[1] * gv[1] + gv[2] * gv[2];
        const double det = h00 * h11 - h01 * h01;
        const double i00 = h11 / det, i01 = -h01 / det, i11 = h00 / det;
        double Ei[3], Ek[3], Pb[3];
        const double eu0 = 1.0 - xc[4 * a] - (gu[0] * C[0] + gu[1] * C[1] + gu[2] * C[2]);
        const double ev0 = 1.0 - yc[4 * a] - (gv[0] * C[0] + gv[1] * C[1] + gv[2] * C[2]);
        for (int m = 0; m < 3; m++) {
            const double ci = i00 * gu[m] + i01 * gv[m];  // column of G^T H^-1 for e_u
            const double ck = i01 * gu[m] + i11 * gv[m];  // for e_v
            Ei[m] = ci;
            Ek[m] = ck;
            Pb[m] = C[m] + ci * eu0 + ck * ev0;
        }
        const double gun = std::sqrt(h00), gvn = std::sqrt(h11);
        for (int e = 0; e < n; e++) {
            const double *E = ell + (size_t)MTR_ELLIPSOID_DOUBLES * e;
            const double *R = E + 6;
            const double ax[3] = {E[3], E[4], E[5]};
            const double rmax = std::max(ax[0], std::max(ax[1], ax[2]));
            // Detector position of the centre and bounding box of the
            // projected bounding sphere.
            const double uc = xc[4 * a] + gu[0] * E[0] + gu[1] * E[1] + gu[2] * E[2] - 1.0;
            const double vc = yc[4 * a] + gv[0] * E[0] + gv[1] * E[1] + gv[2] * E[2] - 1.0;
            const int ilo = std::max(0, (int)std::floor(uc - rmax * gun) - 1);
            const int ihi = std::min(nu - 1, (int)std::ceil(uc + rmax * gun) + 1);
            const int klo = std::max(0, (int)std::floor(vc - rmax * gvn) - 1);
            const int khi = std::min(nv - 1, (int)std::ceil(vc + rmax * gvn) + 1);
            if (ilo > ihi || klo > khi) continue;
            // Ellipsoid frame q = S R (P - centre), S = diag(1/axes).
            double qd[3], qi[3], qk[3], q0[3];
            for (int m = 0; m < 3; m++) {
                const double *Rm = R + 3 * m;
                qd[m] = (Rm[0] * d[0] + Rm[1] * d[1] + Rm[2] * d[2]) / ax[m];
                qi[m] = (Rm[0] * Ei[0] + Rm[1] * Ei[1] + Rm[2] * Ei[2]) / ax[m];
                qk[m] = (Rm[0] * Ek[0] + Rm[1] * Ek[1] + Rm[2] * Ek[2]) / ax[m];
                q0[m] = (Rm[0] * (Pb[0] - E[0]) + Rm[1] * (Pb[1] - E[1]) + Rm[2] * (Pb[2] - E[2])) / ax[m];
            }
            const double A2 = qd[0] * qd[0] + qd[1] * qd[1] + qd[2] * qd[2];
            const double dens = E[15];
            for (int k = klo; k <= khi; k++)
                for (int i = ilo; i <= ihi; i++) {
                    const double x0 = q0[0] + i * qi[0] + k * qk[0];
                    const double x1 = q0[1] + i * qi[1] + k * qk[1];
                    const double x2 = q0[2] + i * qi[2] + k * qk[2];
                    const double B = x0 * qd[0] + x1 * qd[1] + x2 * qd[2];
                    const double Cq = x0 * x0 + x1 * x1 + x2 * x2 - 1.0;
                    const double disc = B * B - A2 * Cq;
                    if (disc > 0) acc[(size_t)k * nu + i] += dens * 2.0 * std::sqrt(disc) / A2;
                }
        }
        for (int64_t j = 0; j < plane; j++) img[j] = (float)acc[(size_t)j];
    });
    return bad ? MTR_ERR_ARG : MTR_OK;
}

extern "C" int mtr_synth_noise(uint64_t seed, int64_t count, float *p, double frac,
                               double *sigma_out)
{
    if (count < 0 || (count > 0 && !p)) return MTR_ERR_ARG;
    double mx = 0.0;
    for (int64_t j = 0; j < count; j++) mx = std::max(mx, (double)p[j]);
    const double sigma = frac * mx;
    Xoshiro256 rng(seed);
    int64_t j = 0;
    while (j < count) {
        double a, b;
        rng.normal2(a, b);
        p[j] = (float)(p[j] + sigma * a);
        if (++j < count) p[j] = (float)(p[j] + sigma * b);
        ++j;
    }
    if (sigma_out) *sigma_out = sigma;
    return MTR_OK;
}

extern "C" int mtr_synth_voxelize(const mtr_volume_size *vol, int n, const double *ell, float *x,
                                  int threads)
{
    if (!vol_ok(vol) || n < 0 || (n > 0 && !ell) || !x) return MTR_ERR_ARG;
    const int64_t sxy = (int64_t)vol->X * vol->Y;
    parallel_for(vol->Z, resolve_threads(threads), [&](int zz) {
        const int Z = vol->Zstart + zz;
        float *pl = x + sxy * zz;
        std::vector<double> acc((size_t)sxy, 0.0);
        for (int e = 0; e < n; e++) {
            const double *E = ell + (size_t)MTR_ELLIPSOID_DOUBLES * e;
            const double *R = E + 6;
            const double rmax = std::max(E[3], std::max(E[4], E[5]));
            if (Z < E[2] - rmax - 1 || Z > E[2] + rmax + 1) continue;
            const int xlo = std::max(vol->Xstart, (int)std::floor(E[0] - rmax));
            const int xhi = std::min(vol->Xend - 1, (int)std::ceil(E[0] + rmax));
            const int ylo = std::max(vol->Ystart, (int)std::floor(E[1] - rmax));
            const int yhi = std::min(vol->Yend - 1, (int)std::ceil(E[1] + rmax));
            const double dz = Z - E[2];
            for (int Y = ylo; Y <= yhi; Y++)
                for (int X = xlo; X <= xhi; X++) {
                    const double dx = X - E[0], dy = Y - E[1];
                    double q = 0;
                    for (int m = 0; m < 3; m++) {
                        const double l = (R[3 * m] * dx + R[3 * m + 1] * dy + R[3 * m + 2] * dz) / E[3 + m];
                        q += l * l;
                    }
                    if (q <= 1.0)
                        acc[(size_t)(Y - vol->Ystart) * vol->X + (X - vol->Xstart)] += E[15];
                }
        }
        for (int64_t j = 0; j < sxy; j++) pl[j] = (float)acc[(size_t)j];
    });
    return MTR_OK;
}

extern "C" int mtr_synth_uniform(uint64_t seed, int64_t n, double *out)
{
    if (n < 0 || (n > 0 && !out)) return MTR_ERR_ARG;
    Xoshiro256 rng(seed);
    for (int64_t i = 0; i < n; i++) out[i] = rng.uniform();
    return MTR_OK;
}

extern "C" int mtr_synth_normal(uint64_t seed, int64_t n, double *out)
{
    if (n < 0 || (n > 0 && !out)) return MTR_ERR_ARG;
    Xoshiro256 rng(seed);
    int64_t i = 0;
    while (i < n) {
        double a, b;
        rng.normal2(a, b);
        out[i++] = a;
        if (i < n) out[i++] = b;
    }
    return MTR_OK;
}
