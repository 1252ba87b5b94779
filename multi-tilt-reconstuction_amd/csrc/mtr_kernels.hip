// HIP kernels of the SIRT / MACE hot path for CDNA4 (gfx950, wave64).
//
// Operator (file:line in Autumn1998/Multi-tilt-Reconstuction):
//   geometry     u = xc0 + xc1 X + xc2 Y + xc3 Z - 1 (FP64)   FunctionOnGPU.cu:12-25
//   FP (splat)   s = sum fl32(w x), c = sum fl32(w)            FunctionOnGPU.cu:27-52, 316-333
//   residual     r = p - (c != 0 ? s/c : s)                    FunctionOnGPU.cu:228-236
//   BP (gather)  x += fl32(S/C) * lambda where C != 0          FunctionOnGPU.cu:54-79, 111-135
//   BPT          x  = S/C of p                                 FunctionOnGPU.cu:82-109
//   z            z = 2 wbar - w                                FunctionOnGPU.cu:268-276
//   Mann         w = rho (2v - z) + (1 - rho) w                MultiAxis.cpp:33-38
//   mean         wbar = sum_a w_a / K                          MultiAxis.cpp:61-66
//
// The forward projection is computed as a ray-driven gather: every
// detector pixel sums the weights of the voxels in its bilinear support,
// so no atomics are involved and the result is deterministic.  Each
// contribution is formed exactly as the reference's splat forms it
// (same FP64 coordinate chain, same weight products, same FP32 rounding);
// only the summation order differs, which the reference leaves undefined.
#include "mtr_kernels.h"

#include <hip/hip_runtime.h>

#include <type_traits>

namespace mtr {

// ------------------------------------------------------------------------
// storage helpers
template <class T>
__device__ __forceinline__ float ldf(const T *p, int64_t i);
template <>
__device__ __forceinline__ float ldf<float>(const float *p, int64_t i) { return p[i]; }
template <>
__device__ __forceinline__ float ldf<__hip_bfloat16>(const __hip_bfloat16 *p, int64_t i)
{
    return __bfloat162float(p[i]);
}
template <>
__device__ __forceinline__ float ldf<__half>(const __half *p, int64_t i) { return __half2float(p[i]); }

template <class T>
__device__ __forceinline__ void stf(T *p, int64_t i, float v);
template <>
__device__ __forceinline__ void stf<float>(float *p, int64_t i, float v) { p[i] = v; }
template <>
__device__ __forceinline__ void stf<__hip_bfloat16>(__hip_bfloat16 *p, int64_t i, float v)
{
    p[i] = __float2bfloat16(v);
}
template <>
__device__ __forceinline__ void stf<__half>(__half *p, int64_t i, float v) { p[i] = __float2half(v); }

// Value of v after a round trip through storage type T.
template <class T>
__device__ __forceinline__ float rnd(float v);
template <>
__device__ __forceinline__ float rnd<float>(float v) { return v; }
template <>
__device__ __forceinline__ float rnd<__hip_bfloat16>(float v) { return __bfloat162float(__float2bfloat16(v)); }
template <>
__device__ __forceinline__ float rnd<__half>(float v) { return __half2float(__float2half(v)); }

// Global (address space 1) view of a pointer held in a struct: lets the
// compiler emit global loads with a scalar base instead of flat loads.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gview(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}

template <class T>
__device__ __forceinline__ float ldg(const __attribute__((address_space(1))) T *p, uint32_t i);
template <>
__device__ __forceinline__ float ldg<float>(const __attribute__((address_space(1))) float *p, uint32_t i)
{
    return p[i];
}
template <>
__device__ __forceinline__ float ldg<__hip_bfloat16>(const __attribute__((address_space(1))) __hip_bfloat16 *p,
                                                     uint32_t i)
{
    const unsigned short b = ((const __attribute__((address_space(1))) unsigned short *)p)[i];
    return __uint_as_float((unsigned)b << 16);
}
template <>
__device__ __forceinline__ float ldg<__half>(const __attribute__((address_space(1))) __half *p, uint32_t i)
{
    const unsigned short b = ((const __attribute__((address_space(1))) unsigned short *)p)[i];
    return __half2float(__ushort_as_half(b));
}

// A pointer that is the same in every lane, moved into scalar registers
// so that loads from it use the scalar-base form with a 32-bit vector
// offset (no per-load 64-bit address arithmetic).
__device__ __forceinline__ const __attribute__((address_space(1))) char *uniform_base(const void *p)
{
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (const __attribute__((address_space(1))) char *)(((uint64_t)hi << 32) | lo);
}

// Element of type T at a byte offset from a scalar base.
template <class T>
__device__ __forceinline__ float ldb(const __attribute__((address_space(1))) char *b, uint32_t off)
{
    return ldg<T>((const __attribute__((address_space(1))) T *)(b + off), 0u);
}

// The reference's coordinate chain, left to right with fused multiply-adds.
__device__ __forceinline__ double coord(const double *c, int X, int Y, int Z)
{
    return fma(c[3], (double)Z, fma(c[2], (double)Y, fma(c[1], (double)X, c[0]))) - 1.0;
}

// Index of the calling wave in its block, as a wave-uniform (scalar) value.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// floor(g) as an int in one instruction (v_cvt_flr_i32_f32).
__device__ __forceinline__ int cvt_flr(float g)
{
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(g));
    return r;
}

// One FP32 tap of a staged detector row: the fraction du = g - floor(g)
// (v_fract_f32, which returns 0x1.fffffep-1 rather than 1.0 for g in
// (-2^-25, 0)) and floor(g) + ib clamped to [0, lmax], as floor(g) clamped
// to [-ib, lmax - ib] through one med3 of g against glo = -ib and
// ghi = lmax - ib + 0.5 (no rounding: only comparisons), returned without
// the + ib, which the caller folds into its row base.
__device__ __forceinline__ int tap_clamped(float g, float glo, float ghi, float &du)
{
    du = __builtin_amdgcn_fractf(g);
    return cvt_flr(__builtin_amdgcn_fmed3f(g, glo, ghi));
}

// Tent weight max(0, 1 - |t|) of a signed distance t (in detector
// columns), formed in FP64 and rounded once, so small weights keep their
// relative precision.
__device__ __forceinline__ float tent(double t) { return (float)fmax(1.0 - fabs(t), 0.0); }

// Select among three per-axis values without a dynamically indexed array.
template <class V>
__device__ __forceinline__ V sel3(V x, V y, V z, int axis)
{
    return axis == 0 ? x : (axis == 1 ? y : z);
}

// Block-wide FP64 sum of one value per thread; result valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < NT / 64; w++) t += sh[w];
    return t;
}

// ------------------------------------------------------------------------
// Forward projection, ray-driven gather.
//
// Block: 64 x 4 pixels of one view of one agent (blockIdx.y).  Thread
// (i, k) walks the planes of its plan and gathers the voxels of its
// support, forming each contribution as fl32(w * x) with w the reference's
// FP64 weight product.
//
// MODE 0: residual r = p - (c != 0 ? s / c : s) with the precomputed
//         normaliser, FP64 sum of r^2 per block into ag.partial.
// MODE 1: raw s and c into s_out / c_out (stateless operator).
// MODE 2: normaliser c into c_out (geometry only, x taken as 1).
struct FpAcc {
    float s, c;
};

template <class T, bool UNIT_X>
__device__ __forceinline__ void fp_take(FpAcc &acc, const T *x, int64_t j, double w)
{
    acc.c += (float)w;
    if (UNIT_X)
        acc.s += (float)w;
    else
        acc.s += (float)(w * (double)ldf<T>(x, j));
}

template <class T, bool UNIT_X>
__device__ FpAcc fp_ray_general(const AgentDev &ag, const ViewCoef &vc, const FpPlan &pl, int i,
                                int k)
{
    FpAcc acc{0.f, 0.f};
    const T *x = reinterpret_cast<const T *>(ag.x);
    const int Dlo = sel3(ag.Xs, ag.Ys, ag.Zs, pl.D), Dhi = sel3(ag.Xe, ag.Ye, ag.Ze, pl.D) - 1;
    const int Plo = sel3(ag.Xs, ag.Ys, ag.Zs, pl.P), Phi = sel3(ag.Xe, ag.Ye, ag.Ze, pl.P) - 1;
    const int Qlo = sel3(ag.Xs, ag.Ys, ag.Zs, pl.Q), Qhi = sel3(ag.Xe, ag.Ye, ag.Ze, pl.Q) - 1;
    const double ps = pl.p0 + i * pl.pi + k * pl.pk;
    const double qs = pl.q0 + i * pl.qi + k * pl.qk;
    // Planes where the support box can meet the volume.
    double dlo = Dlo, dhi = Dhi;
    if (pl.pd != 0.0) {
        double t0 = (Plo - pl.Ep - 1.0 - ps) / pl.pd, t1 = (Phi + pl.Ep + 1.0 - ps) / pl.pd;
        dlo = fmax(dlo, fmin(t0, t1));
        dhi = fmin(dhi, fmax(t0, t1));
    } else if (ps < Plo - pl.Ep - 1.0 || ps > Phi + pl.Ep + 1.0) {
        return acc;
    }
    if (pl.qd != 0.0) {
        double t0 = (Qlo - pl.Eq - 1.0 - qs) / pl.qd, t1 = (Qhi + pl.Eq + 1.0 - qs) / pl.qd;
        dlo = fmax(dlo, fmin(t0, t1));
        dhi = fmin(dhi, fmax(t0, t1));
    } else if (qs < Qlo - pl.Eq - 1.0 || qs > Qhi + pl.Eq + 1.0) {
        return acc;
    }
    if (!(dlo <= dhi)) return acc;
    const int d0 = max(Dlo, (int)floor(dlo)), d1 = min(Dhi, (int)ceil(dhi));
    const int64_t sD = sel3<int64_t>(ag.sx, ag.sy, ag.sz, pl.D), sP = sel3<int64_t>(ag.sx, ag.sy, ag.sz, pl.P),
                  sQ = sel3<int64_t>(ag.sx, ag.sy, ag.sz, pl.Q);
    const int64_t base = -(int64_t)ag.Xs * ag.sx - (int64_t)ag.Ys * ag.sy - (int64_t)ag.Zs * ag.sz;
    const double *cu = vc.c, *cv = vc.c + 4;
    for (int d = d0; d <= d1; d++) {
        const double pc = ps + d * pl.pd, qc = qs + d * pl.qd;
        const int p0 = max(Plo, (int)ceil(pc - pl.Ep - 1e-9));
        const int p1 = min(Phi, (int)floor(pc + pl.Ep + 1e-9));
        const int q0 = max(Qlo, (int)ceil(qc - pl.Eq - 1e-9));
        const int q1 = min(Qhi, (int)floor(qc + pl.Eq + 1e-9));
        for (int p = p0; p <= p1; p++)
            for (int q = q0; q <= q1; q++) {
                const int X = pl.D == 0 ? d : (pl.P == 0 ? p : q);
                const int Y = pl.D == 1 ? d : (pl.P == 1 ? p : q);
                const int Z = pl.D == 2 ? d : (pl.P == 2 ? p : q);
                const double u = coord(cu, X, Y, Z);
                const double v = coord(cv, X, Y, Z);
                const double tu = u - (double)i, tv = v - (double)k;
                if (fabs(tu) < 1.0 && fabs(tv) < 1.0) {
                    // The reference's weight of this tap: (1 - du) when
                    // floor(u) == i, du = u - floor(u) when floor(u) == i - 1.
                    const double wu = tu >= 0.0 ? 1.0 - tu : u - (double)(i - 1);
                    const double wv = tv >= 0.0 ? 1.0 - tv : v - (double)(k - 1);
                    fp_take<T, UNIT_X>(acc, x, base + d * sD + p * sP + q * sQ, wu * wv);
                }
            }
    }
    return acc;
}

// One line of a row-separable ray: the NC lattice points of the support
// starting at the first point inside it, with FP32 tent weights.
template <class T, int NC>
__device__ __forceinline__ void sep_line_load(const T *x, int64_t lb, int64_t sM, double mc, const FpPlan &pl,
                                              double aM, float aMf, int Mlo, unsigned mspan, float (&wv)[NC],
                                              float (&xv)[NC], bool unit_x)
{
    const double mf = floor(mc - pl.Em) + 1.0;
    int m = (int)mf;
    const double t0 = aM * (mf - mc);
#pragma unroll
    for (int j = 0; j < NC; j++) {
        const bool ok = (unsigned)(m - Mlo) <= mspan;
        wv[j] = ok ? tent(fma((double)j, aM, t0)) : 0.0f;
        if (!unit_x) xv[j] = ldf<T>(x, lb + (ok ? m : Mlo) * sM);
        m++;
    }
}

template <class T, bool NEED_C, bool UNIT_X, int NC>
__device__ __forceinline__ void sep_lines(FpAcc &acc, const T *x, int64_t base, int64_t sL, int64_t sM, double ms,
                                          const FpPlan &pl, double aM, float aMf, int Mlo, unsigned mspan, int l0,
                                          int l1)
{
    constexpr int U = 4;  // lines in flight per lane
    int l = l0;
    for (; l + U - 1 <= l1; l += U) {
        float wv[U][NC], xv[U][NC];
#pragma unroll
        for (int q = 0; q < U; q++)
            sep_line_load<T, NC>(x, base + (l + q) * sL, sM, ms + (l + q) * pl.ml, pl, aM, aMf, Mlo, mspan, wv[q],
                                 xv[q], UNIT_X);
#pragma unroll
        for (int q = 0; q < U; q++)
#pragma unroll
            for (int j = 0; j < NC; j++) {
                if (NEED_C) acc.c += wv[q][j];
                acc.s = UNIT_X ? acc.s + wv[q][j] : fmaf(wv[q][j], xv[q][j], acc.s);
            }
    }
    for (; l <= l1; l++) {
        float wv[NC], xv[NC];
        sep_line_load<T, NC>(x, base + l * sL, sM, ms + l * pl.ml, pl, aM, aMf, Mlo, mspan, wv, xv, UNIT_X);
#pragma unroll
        for (int j = 0; j < NC; j++) {
            if (NEED_C) acc.c += wv[j];
            acc.s = UNIT_X ? acc.s + wv[j] : fmaf(wv[j], xv[j], acc.s);
        }
    }
}

template <class T, bool NEED_C, bool UNIT_X>
__device__ void sep_lines_any(FpAcc &acc, const T *x, int64_t base, int64_t sL, int64_t sM, double ms,
                              const FpPlan &pl, double aM, float aMf, int Mlo, unsigned mspan, int l0, int l1)
{
    for (int l = l0; l <= l1; l++) {
        const double mc = ms + l * pl.ml;
        const double mf = floor(mc - pl.Em) + 1.0;
        int m = (int)mf;
        const double t0 = aM * (mf - mc);
        const int64_t lb = base + l * sL;
        for (int j = 0; j < pl.ncand; j++, m++) {
            const bool ok = (unsigned)(m - Mlo) <= mspan;
            const float wgt = ok ? tent(fma((double)j, aM, t0)) : 0.0f;
            if (NEED_C) acc.c += wgt;
            if (UNIT_X) acc.s += wgt;
            else acc.s = fmaf(wgt, ldf<T>(x, lb + (ok ? m : Mlo) * sM), acc.s);
        }
    }
}

// Row-separable views: detector row k is plane W(k) of the volume with
// weight (1 - dv) = 1, and inside the plane u = alpha + aL*l + aM*m.  Each
// line l of the drive axis meets the ray at m*(l); the support
// |u - i| < 1 is the open interval m*(l) +- 1/|aM|, which holds at most
// ncand lattice points, so every lane runs the same fixed-length loop with
// tent weights max(0, 1 - |t|), t = aM (m - m*), in FP32.
template <class T, bool NEED_C, bool UNIT_X>
__device__ FpAcc fp_ray_separable(const AgentDev &ag, const ViewCoef &vc, const FpPlan &pl, int i, int k)
{
    FpAcc acc{0.f, 0.f};
    const T *x = reinterpret_cast<const T *>(ag.x);
    const int w = (int)(k * pl.wsign + pl.woff);
    if (w < sel3(ag.Xs, ag.Ys, ag.Zs, pl.W) || w > sel3(ag.Xe, ag.Ye, ag.Ze, pl.W) - 1) return acc;
    const int Llo = sel3(ag.Xs, ag.Ys, ag.Zs, pl.L), Lhi = sel3(ag.Xe, ag.Ye, ag.Ze, pl.L) - 1;
    const int Mlo = sel3(ag.Xs, ag.Ys, ag.Zs, pl.M), Mhi = sel3(ag.Xe, ag.Ye, ag.Ze, pl.M) - 1;
    const double ms = pl.m0 + i * pl.mi + w * pl.mw;
    double llo = Llo, lhi = Lhi;
    if (pl.ml != 0.0) {
        double t0 = (Mlo - pl.Em - 1.0 - ms) / pl.ml, t1 = (Mhi + pl.Em + 1.0 - ms) / pl.ml;
        llo = fmax(llo, fmin(t0, t1));
        lhi = fmin(lhi, fmax(t0, t1));
    } else if (ms < Mlo - pl.Em - 1.0 || ms > Mhi + pl.Em + 1.0) {
        return acc;
    }
    if (!(llo <= lhi)) return acc;
    const int l0 = max(Llo, (int)floor(llo)), l1 = min(Lhi, (int)ceil(lhi));
    const int64_t sW = sel3<int64_t>(ag.sx, ag.sy, ag.sz, pl.W), sL = sel3<int64_t>(ag.sx, ag.sy, ag.sz, pl.L),
                  sM = sel3<int64_t>(ag.sx, ag.sy, ag.sz, pl.M);
    const int64_t base = -(int64_t)ag.Xs * ag.sx - (int64_t)ag.Ys * ag.sy - (int64_t)ag.Zs * ag.sz + w * sW;
    const double aM = sel3(vc.c[1], vc.c[2], vc.c[3], pl.M);
    const float aMf = (float)aM;
    const unsigned mspan = (unsigned)(Mhi - Mlo);
    switch (pl.ncand) {
        case 2: sep_lines<T, NEED_C, UNIT_X, 2>(acc, x, base, sL, sM, ms, pl, aM, aMf, Mlo, mspan, l0, l1); break;
        case 3: sep_lines<T, NEED_C, UNIT_X, 3>(acc, x, base, sL, sM, ms, pl, aM, aMf, Mlo, mspan, l0, l1); break;
        case 4: sep_lines<T, NEED_C, UNIT_X, 4>(acc, x, base, sL, sM, ms, pl, aM, aMf, Mlo, mspan, l0, l1); break;
        default: sep_lines_any<T, NEED_C, UNIT_X>(acc, x, base, sL, sM, ms, pl, aM, aMf, Mlo, mspan, l0, l1); break;
    }
    return acc;
}

template <class T, int MODE, bool LANES_K>
__global__ __launch_bounds__(256) void k_fp(const AgentDev *__restrict__ ags, float *__restrict__ s_out,
                                           float *__restrict__ c_out)
{
    const AgentDev &ag = ags[blockIdx.y];
    // 64 x 4 pixel tiles (LANES_K: 4 x 64, lanes along detector rows).
    const int TI = LANES_K ? 4 : 64, TK = LANES_K ? 64 : 4;
    const int nti = (ag.nu + TI - 1) / TI, ntk = (ag.nv + TK - 1) / TK;
    int b = blockIdx.x;
    if (b >= nti * ntk * ag.A) return;
    const int bi = b % nti;
    b /= nti;
    const int bk = b % ntk;
    const int a = b / ntk;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int i = bi * TI + (LANES_K ? wid : lane);
    const int k = bk * TK + (LANES_K ? lane : wid);
    const bool active = i < ag.nu && k < ag.nv;
    FpAcc acc{0.f, 0.f};
    if (active) {
        const ViewCoef vc = ag.coef[a];
        const FpPlan pl = ag.plan[a];
        if (pl.kind == 1)
            acc = fp_ray_separable<T, MODE != 0, MODE == 2>(ag, vc, pl, i, k);
        else
            acc = fp_ray_general<T, MODE == 2>(ag, vc, pl, i, k);
    }
    const int64_t n = (int64_t)a * ag.nu * ag.nv + (int64_t)k * ag.pk + (int64_t)i * ag.pi;
    if (MODE == 0) {
        double r2 = 0.0;
        if (active) {
            const float cc = ag.c[n];
            float sv = acc.s;
            if (cc != 0.0f) sv = sv / cc;
            const float rv = ldf<T>(reinterpret_cast<const T *>(ag.p), n) - sv;
            stf<T>(reinterpret_cast<T *>(ag.r), n, rv);
            const double rs = (double)rnd<T>(rv);  // the stored value
            r2 = rs * rs;
        }
        __shared__ double sh[4];
        const double t = block_sum<256>(r2, sh);
        if (threadIdx.x == 0) ag.partial[blockIdx.x] = t;
    } else if (MODE == 1) {
        if (active) {
            s_out[n] = acc.s;
            c_out[n] = acc.c;
        }
    } else {
        if (active) c_out[n] = acc.c;
    }
}

// ------------------------------------------------------------------------
// Forward projection of row-separable agents whose plane axis W is the
// fastest-varying axis of the volume and whose u does not depend on W.
// Block: 4 waves, each G adjacent detector columns i0 .. i0 + G - 1 of one
// view; the 64 lanes are 64 consecutive detector rows k, i.e. 64 adjacent
// planes W(k).  All lanes share the ray paths, so the support is generated
// once per wave, 64 lines at a time: lane q takes line l0 + q, walks the
// union over the G columns of that line's lattice points and keeps the
// voxels with a nonzero weight for any column; a wave prefix sum of the
// per-line counts places them, in (line, point) order, in the wave's LDS
// list as byte offsets and G weights.  Each listed voxel is then one
// coalesced 256-byte load and G FMAs for 64 x G pixels, with the loads of
// the next 8 voxels in flight.  A voxel row that a one-column scheme loads
// once per column it touches (about twice) is loaded once per wave; the
// vector-memory data path is what bounds this kernel.  Every column still
// sums its own taps in (line, point) order, the order of the per-lane path,
// with the same weight expressions, so the sums are those of the per-lane
// path for every G (the extra FMAs add exact zeros).
// Deep boxes (ag.nseg > 1, plane axis X or Y) are cut into Z segments of
// seg_len planes: a block sums its rays' taps inside one segment (the same
// order, restricted to the segment's lines or lattice points) and writes
// them as FP32 partial sums, which k_fp_seg_res adds in segment order and
// turns into the residual; the fused epilogue below serves whole rays.
// The rays in flight on an XCD then share a few MB of voxels instead of
// tens, which its L2 holds.

// Binary digits of n (n >= 0).
__host__ __device__ constexpr int bits_for(int n) { return n < 2 ? 1 : 1 + bits_for(n / 2); }

template <class T, int NC, int G>
__global__ __launch_bounds__(256) void k_fp_stream(const AgentDev *__restrict__ ags)
{
    // lattice points per line and column pair: the columns' first points
    // lie (G - 1) |mi| < 2 (G - 1) apart (|mi| = Em < 2 when NC <= 4), plus
    // one against rounding of the starts
    constexpr int NU = NC + 2 * (G - 1) + 1;
    constexpr int TP = 64 * NU + 8;  // per wave: 64 lines' voxels + 8 padding
    constexpr int CB = bits_for(NU);  // binary digits of a line's count (0 .. NU)
    __shared__ int tap_off[4][TP];
    __shared__ float tap_w[4][G][TP];
    const AgentDev &ag = ags[blockIdx.y];
    const int nti = (ag.nu + 4 * G - 1) / (4 * G), ntk = (ag.nv + 63) / 64, nseg = ag.nseg;
    // XCD-aware order: each XCD takes a contiguous range of (plane group,
    // Z segment, column group and view) so that the rays in flight on one
    // XCD, cut to one segment of seg_len planes, share their voxels in that
    // XCD's L2.  Within a segment the (view, column group) pairs go in the
    // host's seg_order (by where the rays cross the segment's middle
    // plane: all views of one narrow window), else view fastest.
    const int nva = nti * ag.A, nbt = nva * ntk * nseg, per = (nbt + 7) / 8;
    if (blockIdx.x >= (unsigned)(per * 8)) return;  // beyond this agent's partial sums
    int b = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (b >= nbt) {
        // padding block of the 8-way split: its partial sum is still read
        if (threadIdx.x == 0 && nseg == 1) ag.partial[blockIdx.x] = 0.0;
        return;
    }
    int va = b % nva;
    b /= nva;
    const int sg = b % nseg;
    const int bk = b / nseg;
    if (ag.seg_order) va = ag.seg_order[sg * nva + va];
    const int a = va % ag.A;
    const int bi = va / ag.A;
    const int lane = threadIdx.x & 63, wid = wave_id();
    const int i0 = (bi * 4 + wid) * G;
    const int gn = min(G, ag.nu - i0);  // this wave's columns on the detector (may be <= 0)
    const int k = bk * 64 + lane;
    const FpPlan pl = ag.plan[a];
    const ViewCoef vc = ag.coef[a];
    const T *x = reinterpret_cast<const T *>(ag.x);
    const int Wlo = sel3(ag.Xs, ag.Ys, ag.Zs, pl.W), Whi = sel3(ag.Xe, ag.Ye, ag.Ze, pl.W) - 1;
    const int w = k * pl.wsign + (int)pl.woff;
    const bool row_ok = k < ag.nv && w >= Wlo && w <= Whi;
    const int64_t sL = sel3<int64_t>(ag.sx, ag.sy, ag.sz, pl.L), sM = sel3<int64_t>(ag.sx, ag.sy, ag.sz, pl.M),
                  sW = sel3<int64_t>(ag.sx, ag.sy, ag.sz, pl.W);
    const int64_t base0 = -(int64_t)ag.Xs * ag.sx - (int64_t)ag.Ys * ag.sy - (int64_t)ag.Zs * ag.sz;
    // 32-bit byte offsets (the volume is below 4 GiB: the host selects this
    // kernel only then): the lane's plane, and steps along L and M
    const uint32_t lane_off = (uint32_t)(((row_ok ? w : Wlo) - Wlo) * sW) * (uint32_t)sizeof(T);
    const uint32_t bL = (uint32_t)sL * (uint32_t)sizeof(T), bM = (uint32_t)sM * (uint32_t)sizeof(T);
    float acc[G];
#pragma unroll
    for (int g = 0; g < G; g++) acc[g] = 0.f;
    if (gn > 0) {
        const int Llo = sel3(ag.Xs, ag.Ys, ag.Zs, pl.L), Lhi = sel3(ag.Xe, ag.Ye, ag.Ze, pl.L) - 1;
        const int Mlo = sel3(ag.Xs, ag.Ys, ag.Zs, pl.M), Mhi = sel3(ag.Xe, ag.Ye, ag.Ze, pl.M) - 1;
        // this block's segment: planes Z in [zs0, zs1] (the whole box when
        // nseg = 1); whichever of L and M is Z (the host segments only
        // agents whose plane axis W is X or Y) is cut to it, the voxel
        // offsets stay relative to the box's corner
        const int zs0 = nseg > 1 ? ag.Zs + ag.seg_len * sg : ag.Zs;
        const int zs1 = nseg > 1 ? min(ag.Ze - 1, zs0 + ag.seg_len - 1) : ag.Ze - 1;
        const int LloS = pl.L == 2 ? max(Llo, zs0) : Llo, LhiS = pl.L == 2 ? min(Lhi, zs1) : Lhi;
        const int MloS = pl.M == 2 ? max(Mlo, zs0) : Mlo, MhiS = pl.M == 2 ? min(Mhi, zs1) : Mhi;
        double ms[G];  // aW == 0: no plane term
        // lines crossing the segment, united over the wave's columns
        int l0 = LhiS + 1, l1 = LloS - 1;
#pragma unroll
        for (int g = 0; g < G; g++) {
            ms[g] = pl.m0 + (i0 + g) * pl.mi;
            if (g >= gn) continue;
            double llo = LloS, lhi = LhiS;
            bool any = true;
            if (pl.ml != 0.0) {
                double t0 = (MloS - pl.Em - 1.0 - ms[g]) / pl.ml, t1 = (MhiS + pl.Em + 1.0 - ms[g]) / pl.ml;
                llo = fmax(llo, fmin(t0, t1));
                lhi = fmin(lhi, fmax(t0, t1));
            } else if (ms[g] < MloS - pl.Em - 1.0 || ms[g] > MhiS + pl.Em + 1.0) {
                any = false;
            }
            if (any && llo <= lhi) {
                l0 = min(l0, max(LloS, (int)floor(llo)));
                l1 = max(l1, min(LhiS, (int)ceil(lhi)));
            }
        }
        const double aM = sel3(vc.c[1], vc.c[2], vc.c[3], pl.M);
        const unsigned mspan = (unsigned)(MhiS - MloS);
        // scalar base: voxel (Llo, Mlo) of the box's first plane
        const auto xc = uniform_base(x + base0 + (int64_t)Wlo * sW + (int64_t)Llo * sL + (int64_t)Mlo * sM);
        for (int lb = l0; lb <= l1; lb += 64) {
            // this lane's line and its lattice points mb .. mb + NU - 1
            const int l = lb + lane;
            double mc[G], t0[G];
            int mfi[G];
            int mb = 0;
#pragma unroll
            for (int g = 0; g < G; g++) {
                mc[g] = ms[g] + l * pl.ml;
                const double mf = floor(mc[g] - pl.Em) + 1.0;
                mfi[g] = (int)mf;
                t0[g] = aM * (mf - mc[g]);
                mb = g == 0 ? mfi[0] : min(mb, mfi[g]);
            }
            // weight of column g at point m: point m - mf of its line, the
            // same expression as the per-lane path
            float wt[NU][G];
            uint32_t keep = 0;
            int cnt = 0;
#pragma unroll
            for (int j = 0; j < NU; j++) {
                const int m = mb + j;
                const bool ok = l <= l1 && (unsigned)(m - MloS) <= mspan;
                bool any = false;
#pragma unroll
                for (int g = 0; g < G; g++) {
                    const int jg = m - mfi[g];
                    wt[j][g] = ok && g < gn && (unsigned)jg < (unsigned)NC ? tent(fma((double)jg, aM, t0[g])) : 0.0f;
                    any = any || wt[j][g] > 0.0f;
                }
                keep |= any ? 1u << j : 0u;
                cnt += any ? 1 : 0;
            }
            // exclusive prefix sum of the per-line counts over the wave, bit
            // plane by bit plane
            const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
            int pos = 0, total = 0;
#pragma unroll
            for (int q = 0; q < CB; q++) {
                const uint64_t mq = __ballot((cnt >> q) & 1);
                pos += __popcll(mq & below) << q;
                total += __popcll(mq) << q;
            }
            const uint32_t off0 = (uint32_t)(l - Llo) * bL + (uint32_t)(mb - Mlo) * bM;
#pragma unroll
            for (int j = 0; j < NU; j++) {
                if (keep & (1u << j)) {
                    tap_off[wid][pos] = (int)(off0 + (uint32_t)j * bM);
#pragma unroll
                    for (int g = 0; g < G; g++) tap_w[wid][g][pos] = wt[j][g];
                    pos++;
                }
            }
            if (lane < 8) {
                // padding to a multiple of 8: voxel (Llo, Mlo), weight 0
                tap_off[wid][total + lane] = 0;
#pragma unroll
                for (int g = 0; g < G; g++) tap_w[wid][g][total + lane] = 0.0f;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // voxels in groups of 8 held in two register sets: the next
            // group's loads are in flight while the current one is summed
            const int cnt8 = (total + 7) & ~7;
            float xa[8], xb[8];
            auto fetch = [&](int e, float(&xv)[8]) {
#pragma unroll
                for (int q = 0; q < 8; q++) xv[q] = ldb<T>(xc, (uint32_t)tap_off[wid][e + q] + lane_off);
            };
            auto sum = [&](int e, const float(&xv)[8]) {
#pragma unroll
                for (int q = 0; q < 8; q++)
#pragma unroll
                    for (int g = 0; g < G; g++) acc[g] = fmaf(tap_w[wid][g][e + q], xv[q], acc[g]);
            };
            if (total > 0) {
                fetch(0, xa);
                int e = 0;
                for (; e + 16 < cnt8; e += 16) {
                    fetch(e + 8, xb);
                    sum(e, xa);
                    fetch(e + 16, xa);
                    sum(e + 8, xb);
                }
                if (e + 8 < cnt8) {
                    fetch(e + 8, xb);
                    sum(e, xa);
                    sum(e + 8, xb);
                } else {
                    sum(e, xa);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    // rows of planes outside the box see no voxel: sum 0 (their lanes
    // walked plane Wlo in step with the others)
    if (nseg > 1) {
        // the segment's partial sums; k_fp_seg_res adds the segments
        if (k < ag.nv) {
            const int64_t P = (int64_t)ag.A * ag.nu * ag.nv;
#pragma unroll
            for (int g = 0; g < G; g++) {
                if (g >= gn) break;
                const int64_t nidx = (int64_t)a * ag.nu * ag.nv + (int64_t)k * ag.pk + (int64_t)(i0 + g) * ag.pi;
                ag.seg_part[(int64_t)sg * P + nidx] = row_ok ? acc[g] : 0.0f;
            }
        }
        return;
    }
    double r2 = 0.0;
    if (k < ag.nv) {
#pragma unroll
        for (int g = 0; g < G; g++) {
            if (g >= gn) break;
            // with the row-fastest layout (pk = 1) the 64 lanes touch 64
            // consecutive elements of p, c and r
            const int64_t nidx = (int64_t)a * ag.nu * ag.nv + (int64_t)k * ag.pk + (int64_t)(i0 + g) * ag.pi;
            const float cc = ag.c[nidx];
            float sv = row_ok ? acc[g] : 0.0f;
            if (cc != 0.0f) sv = sv / cc;
            const float rv = ldf<T>(reinterpret_cast<const T *>(ag.p), nidx) - sv;
            stf<T>(reinterpret_cast<T *>(ag.r), nidx, rv);
            const double rs = (double)rnd<T>(rv);  // the stored value
            r2 += rs * rs;
        }
    }
    __shared__ double sh[4];
    const double tsum = block_sum<256>(r2, sh);
    if (threadIdx.x == 0) ag.partial[blockIdx.x] = tsum;
}

// ------------------------------------------------------------------------
// Back-projection, voxel-driven gather with LDS-staged projection rows.
//
// Block: 64 (X) x 4 (Y) voxel columns of TZ voxels each; blockIdx.y is the
// agent.  For every view the block stages the detector footprint of its
// tile (rows of the projection, with a one-pixel zero border outside the
// detector) into LDS and every voxel gathers its four taps from there.
// Views whose footprint does not fit the LDS budget gather from HBM.
// BPT: x = S/C (or 0); otherwise x = fl32(x + fl32(S/C) * lambda) where
// C != 0, with S and C accumulated in FP32 over views in order.
template <class T, int TZ, bool BPT>
__global__ __launch_bounds__(256) void k_bp(const AgentDev *__restrict__ ags, float lambda,
                                           const T *__restrict__ src_override, int lds_cap)
{
    extern __shared__ float lds[];
    const AgentDev &ag = ags[blockIdx.y];
    const int ntx = (ag.X + 63) / 64, nty = (ag.Y + 3) / 4, ntz = (ag.Z + TZ - 1) / TZ;
    int b = blockIdx.x;
    if (b >= ntx * nty * ntz) return;
    const int bx = b % ntx;
    b /= ntx;
    const int by = b % nty;
    const int bz = b / nty;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int X0 = ag.Xs + bx * 64, Y0 = ag.Ys + by * 4, Z0 = ag.Zs + bz * TZ;
    const int X = X0 + tx, Y = Y0 + ty;
    const bool active = X < ag.Xe && Y < ag.Ye;
    const int ex = min(64, ag.Xe - X0) - 1, ey = min(4, ag.Ye - Y0) - 1, ez = min(TZ, ag.Ze - Z0) - 1;
    const int nu = ag.nu, nv = ag.nv;
    const int64_t plane = (int64_t)nu * nv;
    // BPT back-projects the projections, SIRT the residual.
    const T *rsrc = src_override ? src_override : reinterpret_cast<const T *>(BPT ? ag.p : ag.r);

    float S[TZ], C[TZ];
#pragma unroll
    for (int t = 0; t < TZ; t++) S[t] = C[t] = 0.f;

    for (int a = 0; a < ag.A; a++) {
        const ViewCoef vc = ag.coef[a];
        const double *cu = vc.c, *cv = vc.c + 4;
        // Footprint of the tile (affine map: extremes at the corners).
        const double u0 = coord(cu, X0, Y0, Z0), v0 = coord(cv, X0, Y0, Z0);
        const double ulo = u0 + fmin(0.0, cu[1] * ex) + fmin(0.0, cu[2] * ey) + fmin(0.0, cu[3] * ez);
        const double uhi = u0 + fmax(0.0, cu[1] * ex) + fmax(0.0, cu[2] * ey) + fmax(0.0, cu[3] * ez);
        const double vlo = v0 + fmin(0.0, cv[1] * ex) + fmin(0.0, cv[2] * ey) + fmin(0.0, cv[3] * ez);
        const double vhi = v0 + fmax(0.0, cv[1] * ex) + fmax(0.0, cv[2] * ey) + fmax(0.0, cv[3] * ez);
        const int fi0 = max((int)floor(ulo) - 1, -1), fi1 = min((int)floor(uhi) + 2, nu);
        const int fk0 = max((int)floor(vlo) - 1, -1), fk1 = min((int)floor(vhi) + 2, nv);
        if (fi1 < 0 || fi0 > nu - 1 || fk1 < 0 || fk0 > nv - 1) continue;  // view misses the tile
        const int W = fi1 - fi0 + 1, H = fk1 - fk0 + 1;
        const bool interior = fi0 >= 0 && fi1 <= nu - 1 && fk0 >= 0 && fk1 <= nv - 1;
        const bool staged = W * H <= lds_cap;
        const T *rv = rsrc + plane * a;
        if (staged) {
            __syncthreads();
            for (int row = ty; row < H; row += 4) {
                const int gk = fk0 + row;
                const bool rok = gk >= 0 && gk < nv;
                for (int col = tx; col < W; col += 64) {
                    const int gi = fi0 + col;
                    float val = 0.f;
                    if (rok && gi >= 0 && gi < nu) val = ldf<T>(rv, (int64_t)gk * nu + gi);
                    lds[row * W + col] = val;
                }
            }
            __syncthreads();
        }
        double u = coord(cu, X, Y, Z0), v = coord(cv, X, Y, Z0);
        if (ag.plan[a].kind == 1) {
            // Row-separable view: v is integral, dv == 0, the k0 + 1 taps
            // carry weight 0 and only the two u-taps of row k0 remain.
#pragma unroll
            for (int t = 0; t < TZ; t++) {
                const double fu = floor(u);
                const int i0 = (int)fu, k0 = (int)v;
                const float du = (float)(u - fu);
                float r00, r01;
                if (staged) {
                    const int li = min(max(i0 - fi0, 0), W - 2), lk = min(max(k0 - fk0, 0), H - 2);
                    const float *q = lds + lk * W + li;
                    r00 = q[0];
                    r01 = q[1];
                } else {
                    const int ia = min(max(i0, 0), nu - 1), ib = min(max(i0 + 1, 0), nu - 1);
                    const int ka = min(max(k0, 0), nv - 1);
                    r00 = ldf<T>(rv, (int64_t)ka * nu + ia);
                    r01 = ldf<T>(rv, (int64_t)ka * nu + ib);
                }
                if (interior) {
                    S[t] += fmaf(du, r01 - r00, r00);
                    C[t] += 1.0f;
                } else {
                    const bool kv = (unsigned)k0 < (unsigned)nv;
                    const float wu0 = kv && (unsigned)i0 < (unsigned)nu ? 1.f - du : 0.f;
                    const float wu1 = kv && (unsigned)(i0 + 1) < (unsigned)nu ? du : 0.f;
                    S[t] += wu0 * r00 + wu1 * r01;
                    C[t] += wu0 + wu1;
                }
                u += cu[3];
                v += cv[3];
            }
            continue;
        }
#pragma unroll
        for (int t = 0; t < TZ; t++) {
            const double fu = floor(u), fv = floor(v);
            const int i0 = (int)fu, k0 = (int)fv;
            const float du = (float)(u - fu), dv = (float)(v - fv);
            float r00, r01, r10, r11;
            if (staged) {
                const int li = min(max(i0 - fi0, 0), W - 2), lk = min(max(k0 - fk0, 0), H - 2);
                const float *q = lds + lk * W + li;
                r00 = q[0];
                r01 = q[1];
                r10 = q[W];
                r11 = q[W + 1];
            } else {
                const int ia = min(max(i0, 0), nu - 1), ib = min(max(i0 + 1, 0), nu - 1);
                const int ka = min(max(k0, 0), nv - 1), kb = min(max(k0 + 1, 0), nv - 1);
                r00 = ldf<T>(rv, (int64_t)ka * nu + ia);
                r01 = ldf<T>(rv, (int64_t)ka * nu + ib);
                r10 = ldf<T>(rv, (int64_t)kb * nu + ia);
                r11 = ldf<T>(rv, (int64_t)kb * nu + ib);
            }
            if (interior) {
                const float top = fmaf(du, r01 - r00, r00);
                const float bot = fmaf(du, r11 - r10, r10);
                S[t] += fmaf(dv, bot - top, top);
                C[t] += 1.0f;
            } else {
                const float wu0 = (unsigned)i0 < (unsigned)nu ? 1.f - du : 0.f;
                const float wu1 = (unsigned)(i0 + 1) < (unsigned)nu ? du : 0.f;
                const float wv0 = (unsigned)k0 < (unsigned)nv ? 1.f - dv : 0.f;
                const float wv1 = (unsigned)(k0 + 1) < (unsigned)nv ? dv : 0.f;
                S[t] += wv0 * (wu0 * r00 + wu1 * r01) + wv1 * (wu0 * r10 + wu1 * r11);
                C[t] += (wu0 + wu1) * (wv0 + wv1);
            }
            u += cu[3];
            v += cv[3];
        }
    }
    if (!active) return;
    T *x = reinterpret_cast<T *>(ag.x);
    const int64_t sxy = ag.sz;
    const int64_t j0 = (int64_t)(X - ag.Xs) * ag.sx + (int64_t)(Y - ag.Ys) * ag.sy + (int64_t)(Z0 - ag.Zs) * sxy;
#pragma unroll
    for (int t = 0; t < TZ; t++) {
        if (t > ez) break;
        const int64_t j = j0 + t * sxy;
        if (BPT) {
            stf<T>(x, j, C[t] != 0.0f ? S[t] / C[t] : 0.0f);
        } else if (C[t] != 0.0f) {
            stf<T>(x, j, fmaf(S[t] / C[t], lambda, ldf<T>(x, j)));
        }
    }
}

// ------------------------------------------------------------------------
// Back-projection of agents whose views are all row-separable with u
// independent of the plane axis W (v = v0 + bW*W, u = u(P1, P2)): the slab
// and axis-Z geometries.  Each thread owns a run of RW voxels along W at a
// fixed (P1, P2); per view it forms u, floor(u) and du once, and the run's
// voxels read detector rows k(W0 + t) of one column pair.  The block's
// detector footprint (RW rows x at most PITCH columns, zero outside the
// detector) is staged in LDS with a fixed pitch, double buffered: the next
// view's rows are fetched into registers while the current view is
// accumulated.  Lanes run along P1 (X, or Y when W = X), waves along P2.
struct BpSepArgs {
    int W, P1, P2;  // axes (0 = X, 1 = Y, 2 = Z)
};

template <int RW, int PITCH, int NT>
struct BpSepStage {
    static constexpr int kElems = RW * PITCH;
    static constexpr int kPer = (kElems + NT - 1) / NT;
};

constexpr int kSepWaves = 8;  // waves along P2 per block (512 threads)

static_assert(sizeof(BpSepView) == 48 && offsetof(BpSepView, k0) == 40, "BpSepView: five doubles, two ints");

// u coefficients c[0..3] of a view through the scalar cache.
__device__ __forceinline__ void load_ucoef(const ViewCoef *tab, int a, double (&c)[4])
{
    const __attribute__((address_space(4))) double *d =
        (const __attribute__((address_space(4))) double *)(tab + a);
#pragma unroll
    for (int q = 0; q < 4; q++) c[q] = d[q];
}

// View constants through the scalar cache (the table is read-only during a
// launch), so that waiting for them never waits for vector loads in flight.
__device__ __forceinline__ BpSepView load_bsv(const BpSepView *tab, int a)
{
    const __attribute__((address_space(4))) double *d =
        (const __attribute__((address_space(4))) double *)(tab + a);
    const __attribute__((address_space(4))) int *k = (const __attribute__((address_space(4))) int *)(d + 5);
    BpSepView v;
    v.c0 = d[0];
    v.a1 = d[1];
    v.a2 = d[2];
    v.umin = d[3];
    v.umax = d[4];
    v.k0 = k[0];
    v.kstep = k[1];
    return v;
}

template <class T, int RW, int PITCH, bool BPT, int VPS>
__global__ __launch_bounds__(64 * kSepWaves) __attribute__((amdgpu_waves_per_eu(6))) void k_bp_sep(
    const AgentDev *__restrict__ ags, float lambda, BpSepArgs sa)
{
    constexpr int NT = 64 * kSepWaves;
    // LDS holds (r[k][i], r[k][i+1] - r[k][i]) pairs so that the interior
    // update is S += r0 + du * d: one 64-bit LDS read and two FP32 ops.
    // VPS views are staged per barrier, in two buffers.
    constexpr int kStage = 2 * VPS * RW * PITCH, kEpi = kSepWaves * 64 * 18 / 2;
    __shared__ float2 lds_raw[kStage > kEpi ? kStage : kEpi];
    float2(*lds)[RW * PITCH] = reinterpret_cast<float2(*)[RW * PITCH]>(lds_raw);
    const AgentDev &ag = ags[blockIdx.y];
    const int n1 = sel3(ag.X, ag.Y, ag.Z, sa.P1), n2 = sel3(ag.X, ag.Y, ag.Z, sa.P2), nw = sel3(ag.X, ag.Y, ag.Z, sa.W);
    const int s1 = sel3(ag.Xs, ag.Ys, ag.Zs, sa.P1), s2 = sel3(ag.Xs, ag.Ys, ag.Zs, sa.P2), sw = sel3(ag.Xs, ag.Ys, ag.Zs, sa.W);
    const int nt1 = (n1 + 63) / 64, nt2 = (n2 + kSepWaves - 1) / kSepWaves, ntw = (nw + RW - 1) / RW;
    const int nb = nt1 * nt2 * ntw;
    // XCD-aware order: consecutive tiles (which share detector rows) run on
    // the same XCD's L2.
    const int per = (nb + 7) / 8;
    const int b = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (blockIdx.x >= (unsigned)(per * 8) || b >= nb) return;
    const int b1 = b % nt1, b2 = (b / nt1) % nt2, bw = b / (nt1 * nt2);
    const int lane = threadIdx.x & 63, wid = wave_id();
    const int P10 = s1 + b1 * 64, P20 = s2 + b2 * kSepWaves, W0 = sw + bw * RW;
    const int p1 = P10 + lane, p2 = P20 + wid;
    const int ew = min(RW, sw + nw - W0) - 1;
    const int nu = ag.nu, nv = ag.nv;
    const int64_t plane = (int64_t)nu * nv;
    const auto src = gview(reinterpret_cast<const T *>(BPT ? ag.p : ag.r));
    const int X = sa.P1 == 0 ? p1 : (sa.P2 == 0 ? p2 : W0);
    const int Y = sa.P1 == 1 ? p1 : (sa.P2 == 1 ? p2 : W0);
    const int Z = sa.P1 == 2 ? p1 : (sa.P2 == 2 ? p2 : W0);
    using St = BpSepStage<RW, PITCH, NT>;

    float S[RW];
#pragma unroll
    for (int t = 0; t < RW; t++) S[t] = 0.f;
    float Cacc = 0.f;  // every row of the run is on the detector (host-checked)
    float2 pre[VPS][St::kPer];
    // this thread's staging slots: fixed (row, column) of the footprint.
    // Consecutive threads take consecutive columns in the [nv][nu] layout
    // and consecutive rows in the row-fastest [nu][nv] layout, so the
    // staging loads are contiguous either way.
    // agent constants used in the view loop, held in registers
    const int pk = ag.pk, pi = ag.pi, nviews = ag.A;
    const BpSepView *const bsv = ag.bsv;
    const bool rowfast = pk == 1;
    int srow[St::kPer], scol[St::kPer];
    // byte offsets of a slot's pair within a footprint whose lowest row and
    // first column are at the origin: rows ascending with W (kstep > 0) or
    // descending (kstep < 0), so that both stay non-negative
    uint32_t oup[St::kPer], odn[St::kPer];
#pragma unroll
    for (int q = 0; q < St::kPer; q++) {
        const int e = threadIdx.x + NT * q;
        const int r = rowfast ? e % RW : e / PITCH, cidx = rowfast ? e / RW : e % PITCH;
        srow[q] = e < St::kElems ? r : RW;  // RW: no slot
        scol[q] = cidx;
        const int rr = e < St::kElems ? r : 0;
        oup[q] = (uint32_t)(rr * pk + cidx * pi) * (uint32_t)sizeof(T);
        odn[q] = (uint32_t)((RW - 1 - rr) * pk + cidx * pi) * (uint32_t)sizeof(T);
    }
    const uint32_t pib = (uint32_t)pi * (uint32_t)sizeof(T);
    const T *const srcp = reinterpret_cast<const T *>(BPT ? ag.p : ag.r);
    // full runs: every staged row is on the detector (host-checked)
    const bool full_run = ew == RW - 1;

    // the tile's footprint window in each view (the host's sep_window),
    // through the scalar cache
    const __attribute__((address_space(4))) int *wtab =
        (const __attribute__((address_space(4))) int *)(ag.sep_win + (int64_t)(b1 + nt1 * b2) * ag.A);
    auto fetch = [&](const BpSepView &bv, int a, float2(&pv)[St::kPer], int &fi0, int &width) {
        fi0 = wtab[2 * a];
        const int fi1 = wtab[2 * a + 1];
        width = fi1 - fi0 + 1;
        const int k0row = bv.k0 + bv.kstep * W0, kstep = bv.kstep;
        const uint32_t vbase = (uint32_t)(plane * a);
        if (full_run && fi0 >= 0 && fi0 + PITCH <= nu - 1) {
            // every slot's pair lies inside the detector: no clamps or masks;
            // a scalar base at the footprint's lowest row and first column
            const bool up = kstep > 0;
            const int krow = up ? k0row : k0row - (RW - 1);
            const auto sp = uniform_base(srcp + (int64_t)plane * a + (int64_t)krow * pk + (int64_t)fi0 * pi);
#pragma unroll
            for (int q = 0; q < St::kPer; q++) {
                const uint32_t o = up ? oup[q] : odn[q];
                const float v0 = ldb<T>(sp, o);
                const float v1 = ldb<T>(sp, o + pib);
                pv[q] = make_float2(v0, v1 - v0);
            }
            return;
        }
#pragma unroll
        for (int q = 0; q < St::kPer; q++) {
            // branch-free: clamp the address into the detector, select 0 outside
            const int gk = k0row + kstep * srow[q], gi = fi0 + scol[q];
            const bool ok = srow[q] <= ew && scol[q] < width;
            const bool ok0 = ok && (unsigned)gi < (unsigned)nu, ok1 = ok && (unsigned)(gi + 1) < (unsigned)nu;
            const uint32_t o = vbase + (uint32_t)min(max(gk, 0), nv - 1) * (uint32_t)pk;
            const float v0 = ldg<T>(src, o + (uint32_t)min(max(gi, 0), nu - 1) * (uint32_t)pi);
            const float v1 = ldg<T>(src, o + (uint32_t)min(max(gi + 1, 0), nu - 1) * (uint32_t)pi);
            const float r0 = ok0 ? v0 : 0.f, r1 = ok1 ? v1 : 0.f;
            pv[q] = make_float2(r0, r1 - r0);
        }
    };
    auto commit = [&](int slot, const float2(&pv)[St::kPer]) {
#pragma unroll
        for (int q = 0; q < St::kPer; q++)
            if (srow[q] < RW) lds[slot][srow[q] * PITCH + scol[q]] = pv[q];
    };
    // one view's taps from its staged footprint
    auto gather = [&](const BpSepView &bv, int cfi0, int cwidth, const float2 *rows) {
        const double u = fma(bv.a1, (double)p1, fma(bv.a2, (double)p2, bv.c0));
        const double fu = floor(u);
        const int i0 = (int)fu;
        const float du = (float)(u - fu);
        const int li = min(max(i0 - cfi0, 0), min(cwidth, PITCH) - 2);
        const float2 *q = rows + li;
        const bool interior = cfi0 >= 0 && cfi0 + cwidth - 1 <= nu - 1;
        // (a view with cwidth < 2 misses this tile: no tap is on the detector)
        if (cwidth >= 2) {
            // S += wA r0 + wB (r1 - r0): wA = 1, wB = du inside the detector
            // (fma by 1 is the plain sum), at its side edges the taps' weights
            // with the off-detector one at 0
            float wA = 1.0f, wB = du;
            if (!interior) {
                const float wu0 = (unsigned)i0 < (unsigned)nu ? 1.f - du : 0.f;
                const float wu1 = (unsigned)(i0 + 1) < (unsigned)nu ? du : 0.f;
                wA = wu0 + wu1;
                wB = wu1;
            }
#pragma unroll
            for (int t = 0; t < RW; t++) {
                const float2 rd = q[t * PITCH];
                S[t] = fmaf(wB, rd.y, fmaf(wA, rd.x, S[t]));
            }
            Cacc += wA;
        }
    };

    // stage s holds views s VPS .. s VPS + VPS - 1 in buffer s & 1; the next
    // stage's loads are in flight while the current one is gathered, views
    // in order
    int fi0[VPS], width[VPS];
    BpSepView bvs[VPS];
#pragma unroll
    for (int v = 0; v < VPS; v++)
        if (v < nviews) {
            bvs[v] = load_bsv(bsv, v);
            fetch(bvs[v], v, pre[v], fi0[v], width[v]);
        }
#pragma unroll
    for (int v = 0; v < VPS; v++)
        if (v < nviews) commit(v, pre[v]);
    __syncthreads();
    for (int a0 = 0; a0 < nviews; a0 += VPS) {
        const int buf = (a0 / VPS) & 1;
        int cfi0[VPS], cwidth[VPS];
        BpSepView cbv[VPS];
#pragma unroll
        for (int v = 0; v < VPS; v++) {
            cfi0[v] = fi0[v];
            cwidth[v] = width[v];
            cbv[v] = bvs[v];
        }
        const int an = a0 + VPS;
#pragma unroll
        for (int v = 0; v < VPS; v++)
            if (an + v < nviews) {
                bvs[v] = load_bsv(bsv, an + v);
                fetch(bvs[v], an + v, pre[v], fi0[v], width[v]);
            }
#pragma unroll
        for (int v = 0; v < VPS; v++)
            if (a0 + v < nviews) gather(cbv[v], cfi0[v], cwidth[v], lds[buf * VPS + v]);
#pragma unroll
        for (int v = 0; v < VPS; v++)
            if (an + v < nviews) commit((buf ^ 1) * VPS + v, pre[v]);
        __syncthreads();
    }
    // Epilogue: each lane holds a run along W (contiguous in memory) at its
    // own P1.  Stored directly, every store instruction would touch 64
    // cache lines 4 bytes at a time; instead the wave transposes its runs
    // through LDS, CH voxels of each run at a time, so that CH lanes write
    // CH consecutive voxels of one run.  Same arithmetic as per-lane.
    constexpr int CH = RW < 16 ? RW : 16, LPI = 64 / CH;  // chunk, lines per instruction
    static_assert(RW % CH == 0 && 64 % CH == 0, "run length must divide into chunks");
    float *ep = reinterpret_cast<float *>(lds_raw) + wid * 64 * 18;  // 64 x 17 values + 64 C
    float *cl = ep + 64 * 17;
    T *x = reinterpret_cast<T *>(ag.x);
    const int64_t sW = sel3<int64_t>(ag.sx, ag.sy, ag.sz, sa.W), sP1 = sel3<int64_t>(ag.sx, ag.sy, ag.sz, sa.P1);
    // voxel index of line 0 of this wave (P1 = P10)
    const int64_t jb = (int64_t)(X - ag.Xs) * ag.sx + (int64_t)(Y - ag.Ys) * ag.sy + (int64_t)(Z - ag.Zs) * ag.sz -
                       (int64_t)lane * sP1;
    const bool wave_ok = p2 < s2 + n2;
    cl[lane] = Cacc;
#pragma unroll
    for (int h = 0; h < RW / CH; h++) {
#pragma unroll
        for (int t = 0; t < CH; t++) ep[lane * 17 + t] = S[CH * h + t];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int tt = lane % CH, tw = CH * h + tt;
#pragma unroll 4
        for (int i = 0; i < CH; i++) {
            const int line = i * LPI + lane / CH;
            const float sv = ep[line * 17 + tt], cv = cl[line];
            if (wave_ok && tw <= ew && P10 + line < s1 + n1) {
                const int64_t j = jb + line * sP1 + tw * sW;
                if (BPT) {
                    stf<T>(x, j, cv != 0.0f ? sv / cv : 0.0f);
                } else if (cv != 0.0f) {
                    stf<T>(x, j, fmaf(sv / cv, lambda, ldf<T>(x, j)));
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// ------------------------------------------------------------------------
// Column plan: every view has v independent of Z and the same v (slab
// geometry with any in-plane axis), so each (X, Y) column lands on the two
// detector rows k0 = floor(v), k0 + 1 with weights (1 - dv, dv) in every
// view, while u runs linearly along the column (step xc3 per plane).  The
// working volume is Z-contiguous.

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// DPP move of a 32-bit value.  dpp_z: lanes the pattern does not reach
// read 0 (for full row masks); dpp_old: they keep `old`.
template <int CTRL>
__device__ __forceinline__ int dpp_z(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, true);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_old(int old, int v)
{
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWS, 0xf, false);
}

// Keys of the segmented scan are biased so that no key is 0, the value a
// lane reads when a DPP pattern points outside the wave or row.
constexpr int kKeyBias = 1 << 24;

// One Hillis-Steele step of a segmented inclusive scan of (a, b): a lane
// adds the value of the lane the DPP pattern CTRL names when both carry
// the same key.  Keys are monotone along the lanes, so equal keys at two
// lanes mean every lane between them shares the key.
template <int CTRL>
__device__ __forceinline__ void seg_step(float &a, float &b, int key)
{
    const bool same = dpp_z<CTRL>(key) == key;
    const float a2 = __int_as_float(dpp_z<CTRL>(__float_as_int(a)));
    const float b2 = __int_as_float(dpp_z<CTRL>(__float_as_int(b)));
    a += same ? a2 : 0.0f;
    b += same ? b2 : 0.0f;
}
// Row-carry step (row_bcast): rows outside ROWS keep their values.
template <int CTRL, int ROWS>
__device__ __forceinline__ void seg_carry(float &a, float &b, int key)
{
    const bool same = dpp_old<CTRL, ROWS>(0, key) == key;
    const float a2 = __int_as_float(dpp_old<CTRL, ROWS>(0, __float_as_int(a)));
    const float b2 = __int_as_float(dpp_old<CTRL, ROWS>(0, __float_as_int(b)));
    a += same ? a2 : 0.0f;
    b += same ? b2 : 0.0f;
}

// Adds the original values of the previous R lanes (wave_shr 1 applied
// r times, across rows) that share the lane's key, in lane order.
template <int R>
__device__ __forceinline__ void seg_linear(float &a, float &b, int key)
{
    float sa = a, sb = b, ra = a, rb = b;
    int rk = key;
#pragma unroll
    for (int r = 0; r < R; r++) {
        rk = dpp_z<0x138>(rk);
        ra = __int_as_float(dpp_z<0x138>(__float_as_int(ra)));
        rb = __int_as_float(dpp_z<0x138>(__float_as_int(rb)));
        const bool same = rk == key;
        sa += same ? ra : 0.0f;
        sb += same ? rb : 0.0f;
    }
    a = sa;
    b = sb;
}

// Segmented inclusive scan over the lanes of a wave, for segments of at
// most 1 + CLS lanes (CLS = 1, 2, 3: seg_linear), at most 8 (CLS = 8) or
// any length (CLS = 64): Hillis-Steele inside each row of 16 lanes
// (row_shr 1, 2, 4, and 8 when CLS = 64) and the row carries (row_bcast 15
// into rows 1 and 3, row_bcast 31 into rows 2 and 3).  The last lane of a
// segment ends with the segment's sums.
template <int CLS>
__device__ __forceinline__ void seg_scan_cls(float &a, float &b, int key)
{
    if constexpr (CLS <= 3) {
        seg_linear<CLS>(a, b, key);
    } else {
        seg_step<0x111>(a, b, key);
        seg_step<0x112>(a, b, key);
        seg_step<0x114>(a, b, key);
        if constexpr (CLS > 8) seg_step<0x118>(a, b, key);
        seg_carry<0x142, 0xa>(a, b, key);
        seg_carry<0x143, 0xc>(a, b, key);
    }
}

// A wave's two detector rows (row k0, row k0 + 1) in LDS, interleaved, by
// slot = detector column + 1.  Either the whole width (msk keeps every
// slot), or a window of W = msk + 1 slots (a power of two) addressed modulo
// W: the strip's columns move monotonically along u for every view, and
// a column's first slot never lies more than jit below the first slot of
// any column before it, so when the next column's slots (at most span past
// its first) would leave the window, the slots below its first slot - jit
// are final: they are written out and cleared.  jit and span come from the
// host for the agent; W > span + jit.
struct ColRows {
    float2 *rw;
    int msk, L, nu;
    bool win;
    int jit, span, lo;
    float *pm, *pn;  // output rows k0 and k0 + 1 (null: off the detector)

    __device__ __forceinline__ float2 &at(int slot) { return rw[slot & msk]; }
    // writes out and clears slots [lo, hi)
    __device__ __forceinline__ void flush(int hi, int lane)
    {
        for (int q = lo + lane; q < hi; q += 64) {
            float2 &c = at(q);
            const float2 v = c;
            const int i = q - 1;
            if (i >= 0 && i < nu) {
                if (pm) pm[i] = v.x;
                if (pn) pn[i] = v.y;
            }
            c = make_float2(0.0f, 0.0f);
        }
        lo = hi > lo ? hi : lo;
    }
    // before a column whose planes have u from fma(c3, z, bxy) - 1 over
    // z in [zs, zs + nz): its first slot from the plane of smallest u
    __device__ __forceinline__ void column(double c3, double bxy, int zs, int nz, int lane)
    {
        if (!win) return;
        const double zmin = c3 >= 0.0 ? (double)zs : (double)(zs + nz - 1);
        const int sa = (int)floor(fma(c3, zmin, bxy) - 1.0) + 1;
        if (sa + span >= lo + msk + 1) {
            flush(min(sa - jit, L), lane);
            wave_lds_sync();
        }
    }
};

// One chunk of 64 planes of one column: lane zi - zc holds plane zi with
// value xv.  Leaders (last lane of each segment) add the segment sums into
// the wave's interleaved rows rw (row k0, row k0 + 1): detector column
// floor(u) gets (1 - du) x, floor(u) + 1 gets du x, in two passes so that
// within one pass every LDS address belongs to one lane; only this wave
// touches these rows, so the summation order is fixed.
template <int CLS, bool TAIL>
__device__ __forceinline__ void col_chunk(float xv, int zi, int nz, double c3, double bxy, int zs, float wm,
                                          float wn, ColRows &rw, int nu, int lane)
{
    const bool ok = !TAIL || zi < nz;
    // the reference's chain: fma(xc3, Z, fma(xc2, Y, fma(xc1, X, xc0))) - 1
    const double u = fma(c3, (double)(zs + zi), bxy) - 1.0;
    const double fl = floor(u);
    const int key = ok ? (int)fl + kKeyBias : 0x40000000 + lane;
    // tap weights rounded from FP64 (no cancellation in 1 - du)
    const double dd = u - fl;
    float ea = (float)(1.0 - dd) * xv, eb = (float)dd * xv;
    seg_scan_cls<CLS>(ea, eb, key);
    // next lane's key (wave_shl 1; lane 63 reads 0)
    const int nkey = dpp_z<0x130>(key);
    const bool lead = ok && nkey != key;
    const int bi = key - kKeyBias + 1;  // slot of detector column floor(u) (column -1 -> slot 0)
    if (lead && (unsigned)bi <= (unsigned)(nu + 1)) {
        float2 q = rw.at(bi);
        q.x = fmaf(wm, ea, q.x);
        q.y = fmaf(wn, ea, q.y);
        rw.at(bi) = q;
    }
    wave_lds_sync();
    if (lead && (unsigned)(bi + 1) <= (unsigned)(nu + 1)) {
        float2 q = rw.at(bi + 1);
        q.x = fmaf(wm, eb, q.x);
        q.y = fmaf(wn, eb, q.y);
        rw.at(bi + 1) = q;
    }
    wave_lds_sync();
}

// The strip's columns in order, chunks of 64 planes read in groups of G
// (G <= Z / 64): the loads of a group are issued together, and the first
// group of the next column is loaded while the current column is scanned,
// so the wave waits on global memory once per column at most.
template <class T, bool UNIT, int CLS, int G>
__device__ __forceinline__ void col_strip(const AgentDev &ag, int j0, int j1, double c0, double c1, double c2,
                                          double c3, ColRows &rw, int lane)
{
    const auto x = gview(reinterpret_cast<const T *>(ag.x));
    const int nz = ag.Z, zs = ag.Zs, nu = ag.nu;
    const int64_t xoff = -(int64_t)ag.Xs * ag.sx - (int64_t)ag.Ys * ag.sy;
    auto column = [&](int j) {
        const int2 xy = ag.col_xy[j];
        return x + (xoff + (int64_t)xy.x * ag.sx + (int64_t)xy.y * ag.sy);
    };
    if (j0 >= j1) return;
    float xn[G > 0 ? G : 1];
    if constexpr (G > 0) {
        const auto xcol = column(j0);
#pragma unroll
        for (int g = 0; g < G; g++) xn[g] = UNIT ? 1.0f : ldg<T>(xcol, (uint32_t)(64 * g + lane));
    }
    for (int j = j0; j < j1; j++) {
        const int2 xy = ag.col_xy[j];
        const float2 cw = ag.col_w[j];
        const double bxy = fma(c2, (double)xy.y, fma(c1, (double)xy.x, c0));
        rw.column(c3, bxy, zs, nz, lane);
        const auto xcol = column(j);
        int zc = 0;
        if constexpr (G > 0) {
            // requires Z >= 64 G: the first group is always whole
            float xg[G];
#pragma unroll
            for (int g = 0; g < G; g++) xg[g] = xn[g];
            if (j + 1 < j1) {
                const auto xnext = column(j + 1);
#pragma unroll
                for (int g = 0; g < G; g++) xn[g] = UNIT ? 1.0f : ldg<T>(xnext, (uint32_t)(64 * g + lane));
            }
            while (true) {
#pragma unroll
                for (int g = 0; g < G; g++)
                    col_chunk<CLS, false>(xg[g], zc + 64 * g + lane, nz, c3, bxy, zs, cw.x, cw.y, rw, nu, lane);
                zc += 64 * G;
                if (zc + 64 * G > nz) break;
#pragma unroll
                for (int g = 0; g < G; g++) xg[g] = UNIT ? 1.0f : ldg<T>(xcol, (uint32_t)(zc + 64 * g + lane));
            }
        }
        for (; zc < nz; zc += 64) {
            const int zi = zc + lane;
            float xv = 0.0f;
            if (zi < nz) xv = UNIT ? 1.0f : ldg<T>(xcol, (uint32_t)zi);
            col_chunk<CLS, true>(xv, zi, nz, c3, bxy, zs, cw.x, cw.y, rw, nu, lane);
        }
    }
}

// Chunks of 64 R planes for views with |xc3| < 1/R (R = 2, 4): lane l
// owns planes R l .. R l + R - 1.  Integer crossings of u are then more
// than R - 1 planes apart, so a lane's planes have at most two keys, and
// when the lane crosses, the planes after the crossing (its tail) share
// the key of the next lane's first plane.  Each lane sums its planes in
// order, hands its tail to the next lane (lane 63 writes its own), and the
// segmented scan runs over lanes keyed by their first plane, with runs
// about R times shorter than over single planes.
template <class T, bool UNIT, int R, int CLS>
__device__ __forceinline__ void col_chunk_r(const float (&xv)[R], int zc, double c3, double bxy, int zs, float wm,
                                            float wn, ColRows &rw, int nu, int lane, int dt)
{
    float A = 0.0f, B = 0.0f, At = 0.0f, Bt = 0.0f;
    int f0 = 0;
    bool tail = false;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const double u = fma(c3, (double)(zs + zc + R * lane + r), bxy) - 1.0;
        const double fl = floor(u);
        const double dd = u - fl;
        const float ea = (float)(1.0 - dd) * xv[r], eb = (float)dd * xv[r];
        if (r == 0) {
            f0 = (int)fl;
            A = ea;
            B = eb;
        } else if ((int)fl == f0) {
            A += ea;
            B += eb;
        } else {
            At += ea;
            Bt += eb;
            tail = true;
        }
    }
    // the previous lane's tail (wave_shr 1; lane 0 reads 0)
    A += __int_as_float(dpp_z<0x138>(__float_as_int(At)));
    B += __int_as_float(dpp_z<0x138>(__float_as_int(Bt)));
    const int key = f0 + kKeyBias;
    seg_scan_cls<CLS>(A, B, key);
    const int nkey = dpp_z<0x130>(key);
    const bool lead = nkey != key;
    const int bi = f0 + 1;  // slot of detector column f0
    const bool own = tail && lane == 63;
    if (lead && (unsigned)bi <= (unsigned)(nu + 1)) {
        float2 q = rw.at(bi);
        q.x = fmaf(wm, A, q.x);
        q.y = fmaf(wn, A, q.y);
        rw.at(bi) = q;
    }
    if (own && (unsigned)(bi + dt) <= (unsigned)(nu + 1)) {
        float2 q = rw.at(bi + dt);
        q.x = fmaf(wm, At, q.x);
        q.y = fmaf(wn, At, q.y);
        rw.at(bi + dt) = q;
    }
    wave_lds_sync();
    if (lead && (unsigned)(bi + 1) <= (unsigned)(nu + 1)) {
        float2 q = rw.at(bi + 1);
        q.x = fmaf(wm, B, q.x);
        q.y = fmaf(wn, B, q.y);
        rw.at(bi + 1) = q;
    }
    if (own && (unsigned)(bi + 1 + dt) <= (unsigned)(nu + 1)) {
        float2 q = rw.at(bi + 1 + dt);
        q.x = fmaf(wm, Bt, q.x);
        q.y = fmaf(wn, Bt, q.y);
        rw.at(bi + 1 + dt) = q;
    }
    wave_lds_sync();
}

// Strip walk with chunks of 64 R planes read in groups of G chunks
// (requires Z >= 64 R G), like col_strip; whole chunks beyond the last
// group are read one at a time, and the planes beyond the last whole chunk
// go through col_chunk with the Hillis-Steele scan.
template <class T, bool UNIT, int R, int G, int CLS>
__device__ __forceinline__ void col_strip_r(const AgentDev &ag, int j0, int j1, double c0, double c1, double c2,
                                            double c3, ColRows &rw, int lane)
{
    const auto x = gview(reinterpret_cast<const T *>(ag.x));
    const int nz = ag.Z, zs = ag.Zs, nu = ag.nu, dt = c3 > 0.0 ? 1 : -1;
    const int64_t xoff = -(int64_t)ag.Xs * ag.sx - (int64_t)ag.Ys * ag.sy;
    auto column = [&](int j) {
        const int2 xy = ag.col_xy[j];
        return x + (xoff + (int64_t)xy.x * ag.sx + (int64_t)xy.y * ag.sy);
    };
    if (j0 >= j1) return;
    float xn[G][R];
    {
        const auto xcol = column(j0);
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int r = 0; r < R; r++) xn[g][r] = UNIT ? 1.0f : ldg<T>(xcol, (uint32_t)(64 * R * g + R * lane + r));
    }
    for (int j = j0; j < j1; j++) {
        const int2 xy = ag.col_xy[j];
        const float2 cw = ag.col_w[j];
        const double bxy = fma(c2, (double)xy.y, fma(c1, (double)xy.x, c0));
        rw.column(c3, bxy, zs, nz, lane);
        const auto xcol = column(j);
        float xg[G][R];
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int r = 0; r < R; r++) xg[g][r] = xn[g][r];
        if (j + 1 < j1) {
            const auto xnext = column(j + 1);
#pragma unroll
            for (int g = 0; g < G; g++)
#pragma unroll
                for (int r = 0; r < R; r++)
                    xn[g][r] = UNIT ? 1.0f : ldg<T>(xnext, (uint32_t)(64 * R * g + R * lane + r));
        }
        int zc = 0;
        while (true) {
#pragma unroll
            for (int g = 0; g < G; g++)
                col_chunk_r<T, UNIT, R, CLS>(xg[g], zc + 64 * R * g, c3, bxy, zs, cw.x, cw.y, rw, nu, lane, dt);
            zc += 64 * R * G;
            if (zc + 64 * R * G > nz) break;
#pragma unroll
            for (int g = 0; g < G; g++)
#pragma unroll
                for (int r = 0; r < R; r++)
                    xg[g][r] = UNIT ? 1.0f : ldg<T>(xcol, (uint32_t)(zc + 64 * R * g + R * lane + r));
        }
        for (; zc + 64 * R <= nz; zc += 64 * R) {
            float xv[R];
#pragma unroll
            for (int r = 0; r < R; r++) xv[r] = UNIT ? 1.0f : ldg<T>(xcol, (uint32_t)(zc + R * lane + r));
            col_chunk_r<T, UNIT, R, CLS>(xv, zc, c3, bxy, zs, cw.x, cw.y, rw, nu, lane, dt);
        }
        for (; zc < nz; zc += 64) {
            const int zi = zc + lane;
            float xv = 0.0f;
            if (zi < nz) xv = UNIT ? 1.0f : ldg<T>(xcol, (uint32_t)zi);
            col_chunk<64, true>(xv, zi, nz, c3, bxy, zs, cw.x, cw.y, rw, nu, lane);
        }
    }
}

// Multi-plane lanes for |xc3| < 1/R with the lane scan class from the
// longest run of lanes sharing a first key, ceil(maxrun / R).
template <class T, bool UNIT, int R, int G>
__device__ __forceinline__ void col_strip_rc(const AgentDev &ag, int j0, int j1, double c0, double c1, double c2,
                                             double c3, ColRows &rw, int lane, int maxrun)
{
    const int m = (maxrun + R - 1) / R;
    if (m <= 2) col_strip_r<T, UNIT, R, G, 1>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
    else if (m == 3) col_strip_r<T, UNIT, R, G, 2>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
    else if (m == 4) col_strip_r<T, UNIT, R, G, 3>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
    else if (m <= 8) col_strip_r<T, UNIT, R, G, 8>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
    else col_strip_r<T, UNIT, R, G, 64>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
}

// One plane per lane: chunk groups of 4, 2 or 1 by the volume depth (none
// below 64 planes: partial chunks only).
template <class T, bool UNIT, int CLS>
__device__ __forceinline__ void col_strip_g(const AgentDev &ag, int j0, int j1, double c0, double c1, double c2,
                                            double c3, ColRows &rw, int lane)
{
    if (ag.Z >= 256) col_strip<T, UNIT, CLS, 4>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
    else if (ag.Z >= 128) col_strip<T, UNIT, CLS, 2>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
    else if (ag.Z >= 64) col_strip<T, UNIT, CLS, 1>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
    else col_strip<T, UNIT, CLS, 0>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
}

// Forward projection: one wave per (strip, view); a strip is the set of
// columns sharing k0.  The wave walks the strip's columns in a fixed
// order; for each column its lanes take consecutive planes (one coalesced
// load per 64 planes), form the taps (1 - du) x into detector column
// floor(u) and du x into floor(u) + 1, and the segmented scan sums the
// lanes that share a detector column.  The last lane of each segment adds
// the two sums, weighted (1 - dv) and dv, into the wave's two LDS detector
// rows.  The view's plane step |xc3| fixes, once per wave, the planes per
// lane (1, 2 or 4) and the scan depth; the volume depth fixes how many
// chunks are read at a time.  The rows are written out as partial rows pm
// (row k0) and pn (row k0 + 1); k_fp_col_res adds them.  UNIT: x taken as
// 1 (the normaliser).
constexpr int kColWaves = 4;

template <class T, bool UNIT>
__global__ __launch_bounds__(64 * kColWaves) void k_fp_col(const AgentDev *__restrict__ ags)
{
    extern __shared__ float col_rows[];
    const AgentDev &ag = ags[blockIdx.y];
    const int nvg = (ag.A + kColWaves - 1) / kColWaves;
    // XCD-aware order: strips are dealt to the 8 XCDs in turn (strip
    // lengths vary smoothly across the volume, so contiguous ranges would
    // load the XCDs unevenly), and the view groups of one strip run on one
    // XCD back to back, sharing the strip's columns in its L2
    const int spx = (ag.nstrip + 7) / 8;  // strips per XCD
    const int idx = blockIdx.x >> 3, sx = idx / nvg;
    if (sx >= spx) return;
    const int s = sx * 8 + (blockIdx.x & 7);
    if (s >= ag.nstrip) return;
    const int lane = threadIdx.x & 63, wid = wave_id();
    const int a = (idx - sx * nvg) * kColWaves + wid;
    if (a >= ag.A) return;  // no block-wide barrier below
    const int nu = ag.nu, nv = ag.nv, L = nu + 2;
    const int k0 = s - 1;
    const int64_t row = ((int64_t)a * nv + k0) * nu;
    // the wave's two detector rows, interleaved: (row k0, row k0 + 1), the
    // whole width or a window of col_win slots
    const int nslot = ag.col_win > 0 ? ag.col_win : L;
    ColRows rw;
    rw.rw = reinterpret_cast<float2 *>(col_rows) + wid * nslot;
    rw.win = ag.col_win > 0;
    rw.msk = rw.win ? ag.col_win - 1 : 0x7fffffff;
    rw.L = L;
    rw.nu = nu;
    rw.jit = ag.col_jit;
    rw.span = ag.col_span;
    rw.lo = 0;
    rw.pm = k0 >= 0 ? ag.pm + row : nullptr;
    rw.pn = k0 + 1 < nv ? ag.pn + row + nu : nullptr;
    for (int q = lane; q < nslot; q += 64) rw.rw[q] = make_float2(0.0f, 0.0f);
    wave_lds_sync();
    const ViewCoef vc = ag.coef[a];
    const double c0 = vc.c[0], c1 = vc.c[1], c2 = vc.c[2], c3 = vc.c[3];
    // longest run of consecutive planes sharing floor(u): 1/|xc3| + 1
    const double ac3 = fabs(c3);
    const int maxrun = ac3 > 0.0 ? (int)fmin(64.0, floor(1.0 / ac3) + 1.0) : 64;
    const int j0 = ag.strip[s], j1 = ag.strip[s + 1];
    // lanes own R planes where |xc3| < 1/R; about 256 planes are read per group
    if (ac3 < 0.999 / 4 && ag.Z >= 512) {
        col_strip_rc<T, UNIT, 4, 2>(ag, j0, j1, c0, c1, c2, c3, rw, lane, maxrun);
    } else if (ac3 < 0.999 / 4 && ag.Z >= 256) {
        col_strip_rc<T, UNIT, 4, 1>(ag, j0, j1, c0, c1, c2, c3, rw, lane, maxrun);
    } else if (ac3 < 0.999 / 2 && ag.Z >= 256) {
        col_strip_rc<T, UNIT, 2, 2>(ag, j0, j1, c0, c1, c2, c3, rw, lane, maxrun);
    } else if (ac3 < 0.999 / 2 && ag.Z >= 128) {
        col_strip_rc<T, UNIT, 2, 1>(ag, j0, j1, c0, c1, c2, c3, rw, lane, maxrun);
    } else if (maxrun <= 2) col_strip_g<T, UNIT, 1>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
    else if (maxrun == 3) col_strip_g<T, UNIT, 2>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
    else if (maxrun == 4) col_strip_g<T, UNIT, 3>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
    else if (maxrun <= 8) col_strip_g<T, UNIT, 8>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
    else col_strip_g<T, UNIT, 64>(ag, j0, j1, c0, c1, c2, c3, rw, lane);
    wave_lds_sync();
    rw.flush(L, lane);  // the rest of the rows (all of them without a window)
}

// ------------------------------------------------------------------------
// Column-plan forward projection by prefix moments.
//
// Along a column (X, Y) the planes z = 0 .. Z-1 of the box (absolute Zs + z)
// project in view a to u(z) = ub + c3 z (ub = fma(c3, Zs, bxy) - 1, bxy the
// reference's chain), all onto rows k0, k0 + 1.  Detector column i collects
//     t(i) = sum_z x(z) max(0, 1 - |u(z) - i|),
// the bilinear tap weights of every plane that reaches it.  With the
// column's prefix sums P0(b) = sum_{z<b} x(z) and P1(b) = sum_{z<b} z x(z),
// the ramp moment X2(s) = sum_{z<s} x(z) (s - z) = s P0(ceil s) - P1(ceil s)
// is exact for any real s (clamping ceil s to [0, Z]), and the tent sum is
// its second difference at the plane positions where u crosses i - 1, i and
// i + 1:
//     t(i) = |c3| (X2(s_{i+1}) - 2 X2(s_i) + X2(s_{i-1})),  s_m = (m - ub) / c3,
// for either sign of c3 (the ramp is continuous, so a plane at a crossing
// adds the same whichever side it is counted on).  A column then costs one
// pass per 62 detector columns it reaches (|c3| Z + 2 of them) instead of a
// tap and a segmented reduction per plane: lane l forms X2 at s_{p-1+l},
// lanes l - 1 and l - 2 hand theirs over (DPP shifts), and lanes 2..63 own
// pixels p .. p + 61.  The moments are formed in FP64 (the second
// difference cancels about 16 of its 53 bits) once per column and block,
// shared by the block's eight views (one per wave; the host orders the views
// by the pixels a column reaches, so a block's views cost alike); t is rounded once to
// FP32 and added, weighted (1 - dv, dv), into the wave's two detector rows
// in the strip's fixed column order, with the row window of ColRows.  An
// untilted view (c3 = 0) puts every plane at ub: (1 - du) P0(Z) and
// du P0(Z) at floor(ub) and floor(ub) + 1.  UNIT (the normaliser, x = 1)
// takes the moments in closed form, P0 = b, P1 = b (b - 1) / 2.
constexpr int kPfxWaves = 8;
constexpr int kPfxMaxCols = 8;
constexpr int kPfxLds = 65536;

// 64-bit DPP moves (both halves with the same pattern): zero-fill and
// row-masked forms.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)dpp_z<CTRL>((int)(unsigned)b), hi = (unsigned)dpp_z<CTRL>((int)(unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_d_rows(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)dpp_old<CTRL, ROWS>(0, (int)(unsigned)b);
    const unsigned hi = (unsigned)dpp_old<CTRL, ROWS>(0, (int)(unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double readlane_d(double v, int l)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// Inclusive sum over the wave: Hillis-Steele inside rows of 16 lanes, then
// the row carries (row_bcast 15 into rows 1 and 3, row_bcast 31 into 2, 3).
__device__ __forceinline__ double wave_scan_d(double v)
{
    v += dpp_d<0x111>(v);
    v += dpp_d<0x112>(v);
    v += dpp_d<0x114>(v);
    v += dpp_d<0x118>(v);
    v += dpp_d_rows<0x142, 0xa>(v);
    v += dpp_d_rows<0x143, 0xc>(v);
    return v;
}

// Row slots per wave: the window, or whole rows with room past their end
// for a pass that runs beyond the last pixel (its slots there only ever
// receive zeros and are never written out).
__host__ __device__ inline int pfx_slots(int col_win, int L) { return col_win > 0 ? col_win : ((L + 1) & ~1) + 128; }

// Columns whose moments one block stages at a time (0: does not fit).
__host__ __device__ inline int pfx_cols(int nslot, int nz)
{
    const int left = kPfxLds - kPfxWaves * nslot * (int)sizeof(float2);
    const int per = (nz + 1) * 16;
    if (left < per) return 0;
    return left / per < kPfxMaxCols ? left / per : kPfxMaxCols;
}
__host__ __device__ inline int pfx_lds_bytes(int nslot, int nz, bool unit)
{
    return kPfxWaves * nslot * (int)sizeof(float2) + (unit ? 0 : pfx_cols(nslot, nz) * (nz + 1) * 16);
}

// The first 256 planes of a column, 4 per lane, loaded ahead of its
// moments (the loads of the next round overlap this round's evaluation).
template <class T>
__device__ __forceinline__ void pfx_load(const __attribute__((address_space(1))) T *xcol, int zb, int nz, int lane,
                                         float (&v)[4])
{
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int z = zb + 4 * lane + r;
        v[r] = z < nz ? ldg<T>(xcol, (uint32_t)z) : 0.0f;
    }
}

// Prefix moments of one column (the calling wave): m[b] = (P0(b), P1(b))
// for b = 0 .. Z, 4 planes per lane and 256 per pass; `first` holds the
// first pass's values.
template <class T>
__device__ __forceinline__ void pfx_column(const __attribute__((address_space(1))) T *xcol, int nz, double2 *m,
                                           int lane, const float (&first)[4])
{
    double car0 = 0.0, car1 = 0.0;
    float v[4] = {first[0], first[1], first[2], first[3]};
    for (int zb = 0; zb < nz; zb += 256) {
        if (zb > 0) pfx_load<T>(xcol, zb, nz, lane, v);
        double q0[4], q1[4], t0 = 0.0, t1 = 0.0;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const double e = (double)v[r];
            q0[r] = t0;
            q1[r] = t1;
            t0 += e;
            t1 = fma((double)(zb + 4 * lane + r), e, t1);
        }
        const double i0 = wave_scan_d(t0), i1 = wave_scan_d(t1);
        const double x0 = dpp_d<0x138>(i0) + car0, x1 = dpp_d<0x138>(i1) + car1;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int z = zb + 4 * lane + r;
            if (z < nz) m[z] = make_double2(x0 + q0[r], x1 + q1[r]);
        }
        car0 += readlane_d(i0, 63);
        car1 += readlane_d(i1, 63);
    }
    if (lane == 0) m[nz] = make_double2(car0, car1);
}

// Column table entries through the scalar cache (uniform index).
__device__ __forceinline__ int2 ucol_xy(const int2 *tab, int j)
{
    const __attribute__((address_space(4))) int *t = (const __attribute__((address_space(4))) int *)tab;
    const int q = 2 * __builtin_amdgcn_readfirstlane(j);
    return make_int2(t[q], t[q + 1]);
}
__device__ __forceinline__ float2 ucol_w(const float2 *tab, int j)
{
    const __attribute__((address_space(4))) float *t = (const __attribute__((address_space(4))) float *)tab;
    const int q = 2 * __builtin_amdgcn_readfirstlane(j);
    return make_float2(t[q], t[q + 1]);
}

__device__ __forceinline__ float readlane_f(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// One column in one view (c3 != 0), R = 1 or 2 boundaries per lane: a
// pass covers boundaries m = p - 1 + R lane + r and pixels i = m (lane 0's
// first and lane 63's last pixel lack a neighbour and add zero), 64 R - 2
// pixels.  Every pixel slot of the pass is read and written back; pixels
// outside the column's on-detector range [ilo, ihi] add exactly zero.
// 64 R <= window slots, so no two pixels of a pass share a slot.  With two
// per lane the passes start at even pixels and a lane's two slots are one
// aligned 16-byte row access (consecutive lanes on consecutive addresses).
template <int R, bool UNIT, bool WIN>
__device__ __forceinline__ void pfx_eval(ColRows &rw, const double2 *m, int nz, double ub, int ilo, int ihi,
                                         double inv, float wa, float wb, int lane)
{
    constexpr int PP = 64 * R - 2;
    const double sb = -ub * inv;
    const int p0 = R == 2 ? (ilo & ~1) : ilo;
    // boundary of the lane's first slot, m = p - 1 + R lane, kept as an
    // exact FP64 integer across passes
    double mpos = (double)(p0 - 1 + R * lane);
    // slots outside [ilo, ihi] add zero, and so do lane 0's first and
    // lane 63's last slot
    const unsigned span = (unsigned)(ihi - ilo);
    const int rel = R * lane - 1 - ilo;
    const bool edge_lo = lane == 0, edge_hi = lane == 63;
    for (int p = p0; p <= ihi; p += PP, mpos += (double)PP) {
        const double s0 = fma(mpos, inv, sb);
        double X2[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            // X2(s) = s P0(b) - P1(b) for any b in [ceil(s), floor(s) + 1]
            // (a plane at z = s has weight 0): b = floor(s) + 1 from the
            // FP64 s, clamped to [1, Z]; s < 0 is taken as 0 (X2 = 0)
            const double sg = fmax(r == 0 ? s0 : s0 + inv, 0.0);
            const int b = min((int)sg, nz - 1) + 1;
            double2 mm;
            if constexpr (UNIT) {
                mm = make_double2((double)b, 0.5 * (double)b * (double)(b - 1));
            } else {
                mm = m[b];
            }
            X2[r] = fma(sg, mm.x, -mm.y);
        }
        const double left = dpp_d<0x138>(X2[R - 1]);  // lane - 1's last (wave_shr 1)
        const double right = dpp_d<0x130>(X2[0]);     // lane + 1's first (wave_shl 1)
        float t[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const double xm = r == 0 ? left : X2[r - 1];
            const double xp = r == R - 1 ? right : X2[r + 1];
            const float tv = (float)(fma(-2.0, X2[r], xm) + xp);
            bool zero = (unsigned)(p + rel + r) > span;
            if (r == 0) zero = zero || edge_lo;
            if (r == R - 1) zero = zero || edge_hi;
            t[r] = zero ? 0.0f : tv;
        }
        const int sl0 = p + R * lane;
        const int g = WIN ? (sl0 & rw.msk) : sl0;
        if constexpr (R == 1) {
            const float2 rv = rw.rw[g];
            rw.rw[g] = make_float2(fmaf(wa, t[0], rv.x), fmaf(wb, t[0], rv.y));
        } else {
            float4 *q = reinterpret_cast<float4 *>(rw.rw + g);
            const float4 rv = *q;
            *q = make_float4(fmaf(wa, t[0], rv.x), fmaf(wb, t[0], rv.y), fmaf(wa, t[1], rv.z), fmaf(wb, t[1], rv.w));
        }
    }
}

template <class T, bool UNIT, bool WIN>
__global__ __launch_bounds__(64 * kPfxWaves) void k_fp_pfx(const AgentDev *__restrict__ ags)
{
    extern __shared__ float4 pfx_raw[];
    const AgentDev &ag = ags[blockIdx.y];
    const int nvg = (ag.A + kPfxWaves - 1) / kPfxWaves;
    // XCD-aware order as in k_fp_col: strips dealt to the XCDs in turn, the
    // view groups of one strip back to back on one XCD
    const int spx = (ag.nstrip + 7) / 8;
    const int idx = blockIdx.x >> 3, sx = idx / nvg;
    if (sx >= spx) return;
    const int s = sx * 8 + (blockIdx.x & 7);
    if (s >= ag.nstrip) return;
    const int lane = threadIdx.x & 63, wid = wave_id();
    // wave slot g of the strip's view groups takes the view the host put
    // there (views of alike cost share a block), read through the scalar cache
    const int g = (idx - sx * nvg) * kPfxWaves + wid;
    const bool view = g < ag.A;
    const int a = view ? ((const __attribute__((address_space(4))) FpPlan *)ag.plan)[g].pfx_view : 0;
    if (UNIT && !view) return;  // no block-wide barrier without staged moments
    const int nu = ag.nu, nv = ag.nv, L = nu + 2, nz = ag.Z, zs = ag.Zs;
    const int nslot = pfx_slots(WIN ? ag.col_win : 0, L);
    const int ncol = __builtin_amdgcn_readfirstlane(UNIT ? kPfxMaxCols : pfx_cols(nslot, nz));
    float2 *rows = reinterpret_cast<float2 *>(pfx_raw);
    double2 *mom = reinterpret_cast<double2 *>(rows + kPfxWaves * nslot);
    const int k0 = s - 1;
    ColRows rw;
    rw.rw = rows + wid * nslot;
    rw.win = WIN;
    rw.msk = WIN ? ag.col_win - 1 : 0x7fffffff;
    rw.L = L;
    rw.nu = nu;
    rw.jit = ag.col_jit;
    rw.span = ag.col_span;
    rw.lo = 0;
    const int64_t row = ((int64_t)(view ? a : 0) * nv + k0) * nu;
    rw.pm = view && k0 >= 0 ? ag.pm + row : nullptr;
    rw.pn = view && k0 + 1 < nv ? ag.pn + row + nu : nullptr;
    for (int q = lane; q < nslot; q += 64) rw.rw[q] = make_float2(0.0f, 0.0f);
    wave_lds_sync();
    double uc[4];
    load_ucoef(ag.coef, __builtin_amdgcn_readfirstlane(view ? a : 0), uc);
    const double c0 = uc[0], c1 = uc[1], c2 = uc[2], c3 = uc[3];
    const double inv = c3 != 0.0 ? 1.0 / c3 : 0.0;
    const float ac3 = (float)fabs(c3);
    // boundaries per lane: one where a column reaches at most 62 pixels
    // (|c3| (Z - 1) + 3 at most), two otherwise (64 R slots in the window)
    const int npix = (int)(fabs(c3) * (double)(nz - 1)) + 3;
    const int R = __builtin_amdgcn_readfirstlane(npix <= 62 || (WIN && nslot < 128) ? 1 : 2);
    const int j0 = __builtin_amdgcn_readfirstlane(ag.strip[s]), j1 = __builtin_amdgcn_readfirstlane(ag.strip[s + 1]);
    const auto x = gview(reinterpret_cast<const T *>(ag.x));
    const int64_t xoff = -(int64_t)ag.Xs * ag.sx - (int64_t)ag.Ys * ag.sy;
    auto column = [&](int2 xy) { return x + (xoff + (int64_t)xy.x * ag.sx + (int64_t)xy.y * ag.sy); };
    const int zp = nz + 1;
    // lane q holds the table entry of the round's column q: loaded a round
    // ahead, with the first 256 planes of the column the wave stages next
    int2 txy = make_int2(0, 0);
    float2 tw = make_float2(0.0f, 0.0f);
    float xpre[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (lane < min(ncol, j1 - j0)) {
        txy = ag.col_xy[j0 + lane];
        tw = ag.col_w[j0 + lane];
    }
    if (!UNIT && wid < min(ncol, j1 - j0)) pfx_load<T>(column(ucol_xy(ag.col_xy, j0 + wid)), 0, nz, lane, xpre);
    for (int jr = j0; jr < j1; jr += ncol) {
        const int nc = __builtin_amdgcn_readfirstlane(min(ncol, j1 - jr)), jn = jr + ncol,
                  ncn = __builtin_amdgcn_readfirstlane(max(min(ncol, j1 - jn), 0));
        if constexpr (!UNIT) {
            if (wid < nc) pfx_column<T>(column(ucol_xy(ag.col_xy, jr + wid)), nz, mom + wid * zp, lane, xpre);
            __syncthreads();
        }
        // per-column constants of this view, lane q for column jr + q, with
        // the reference's chain (and the host's, for the window bounds)
        const double bxy = fma(c2, (double)txy.y, fma(c1, (double)txy.x, c0));
        const double ubq = fma(c3, (double)zs, bxy) - 1.0;
        const double ueq = fma(c3, (double)(zs + nz - 1), bxy) - 1.0;
        const int flo = (int)floor(fmin(ubq, ueq)), fhi = (int)floor(fmax(ubq, ueq));
        const float waq = tw.x * ac3, wbq = tw.y * ac3, wzq = tw.x, wyq = tw.y;
        // the next round's table entries and planes load during this round
        if (lane < ncn) {
            txy = ag.col_xy[jn + lane];
            tw = ag.col_w[jn + lane];
        }
        if (!UNIT && wid < ncn) pfx_load<T>(column(ucol_xy(ag.col_xy, jn + wid)), 0, nz, lane, xpre);
        if (view) {
            for (int q = 0; q < nc; q++) {
                const double ub = readlane_d(ubq, q);
                const int lo = __builtin_amdgcn_readlane(flo, q), hi = __builtin_amdgcn_readlane(fhi, q);
                const double2 *m = mom + q * zp;
                if constexpr (WIN) {
                    // slots below the column's first slot less the jitter bound
                    // are final once its slots would leave the window
                    const int sa = lo + 1;
                    if (sa + rw.span >= rw.lo + rw.msk + 1) {
                        rw.flush(min(sa - rw.jit, L), lane);
                        wave_lds_sync();
                    }
                }
                if (c3 == 0.0) {
                    if (lane < 2) {
                        const double fl = floor(ub), du = ub - fl;
                        const int i = (int)fl + lane;
                        const double S = UNIT ? (double)nz : m[nz].x;
                        const float t = (float)((lane == 0 ? 1.0 - du : du) * S);
                        if (i >= 0 && i < nu) {
                            float2 &c = rw.at(i + 1);
                            c.x = fmaf(readlane_f(wzq, q), t, c.x);
                            c.y = fmaf(readlane_f(wyq, q), t, c.y);
                        }
                    }
                } else {
                    const int ilo = max(lo, 0), ihi = min(hi + 1, nu - 1);
                    const float wa = readlane_f(waq, q), wb = readlane_f(wbq, q);
                    if (R == 1) pfx_eval<1, UNIT, WIN>(rw, m, nz, ub, ilo, ihi, inv, wa, wb, lane);
                    else pfx_eval<2, UNIT, WIN>(rw, m, nz, ub, ilo, ihi, inv, wa, wb, lane);
                }
                wave_lds_sync();
            }
        }
        if constexpr (!UNIT) __syncthreads();
    }
    wave_lds_sync();
    if (view) rw.flush(L, lane);
}

// s = pm + pn; MODE 0: residual r = p - (c != 0 ? s/c : s) with FP64
// per-block sums of r^2; MODE 2: the normaliser c = s into c_out.
template <class T, int MODE>
__global__ __launch_bounds__(256) void k_fp_col_res(const AgentDev *__restrict__ ags, float *__restrict__ c_out)
{
    const AgentDev &ag = ags[blockIdx.y];
    const int64_t P = (int64_t)ag.A * ag.nu * ag.nv;
    double r2 = 0.0;
    for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < P; n += (int64_t)gridDim.x * 256) {
        const float sv0 = ag.pm[n] + ag.pn[n];
        if (MODE == 2) {
            c_out[n] = sv0;
        } else {
            const float cc = ag.c[n];
            float sv = sv0;
            if (cc != 0.0f) sv = sv / cc;
            const float rv = ldf<T>(reinterpret_cast<const T *>(ag.p), n) - sv;
            stf<T>(reinterpret_cast<T *>(ag.r), n, rv);
            const double rs = (double)rnd<T>(rv);
            r2 += rs * rs;
        }
    }
    if (MODE == 0) {
        __shared__ double sh[4];
        const double t = block_sum<256>(r2, sh);
        if (threadIdx.x == 0) ag.partial[blockIdx.x] = t;
    }
}

// Segmented streaming forward projection: s = the segments' partial sums
// added in segment order, then the residual r = p - (c != 0 ? s/c : s) and
// FP64 per-block sums of r^2, as the fused epilogue forms them.
template <class T>
__global__ __launch_bounds__(256) void k_fp_seg_res(const AgentDev *__restrict__ ags)
{
    const AgentDev &ag = ags[blockIdx.y];
    const int64_t P = (int64_t)ag.A * ag.nu * ag.nv;
    double r2 = 0.0;
    for (int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x; n < P; n += (int64_t)gridDim.x * 256) {
        float sv = ag.seg_part[n];
        for (int q = 1; q < ag.nseg; q++) sv += ag.seg_part[(int64_t)q * P + n];
        const float cc = ag.c[n];
        if (cc != 0.0f) sv = sv / cc;
        const float rv = ldf<T>(reinterpret_cast<const T *>(ag.p), n) - sv;
        stf<T>(reinterpret_cast<T *>(ag.r), n, rv);
        const double rs = (double)rnd<T>(rv);
        r2 += rs * rs;
    }
    __shared__ double sh[4];
    const double t = block_sum<256>(r2, sh);
    if (threadIdx.x == 0) ag.partial[blockIdx.x] = t;
}

// Back-projection of column-plan agents.  Block: 64 lanes along P1 (X, or
// Y when col_p1y) x 8 waves along the other axis; each thread owns a run
// of RW voxels along Z.  All voxels of a column read rows k0, k0 + 1 with
// weights (1 - dv, dv) in every view, so the tile's rows are fixed for the
// launch; per view the block stages them over the tile's u range as
// (r[i], r[i+1] - r[i]) pairs (zero outside the detector, double
// buffered), and voxel t of the run gathers both rows at u0 + t xc3.
// The normaliser C of a voxel is geometry only: (wk0 + wk1) times the
// summed u weights, one per view less the weight of taps that fall off the
// detector's sides.  MODE 2 computes it once into ag.bpc; MODE 0 (SIRT
// update) and MODE 1 (BPT) read it.
constexpr int kColBpWaves = 8;
constexpr int kColElems = 7 * 64 * kColBpWaves;  // staged pairs per view (rows x pitch)

int col_elems() { return kColElems; }

// Runs of the column back-projection start at absolute multiples of rw, so
// a box [Zs, Ze) is covered by the runs from col_run_base(Zs, rw) on; a box
// that straddles a multiple of rw needs one run more than ceil(Z / rw).
// Shared by the kernel and the host's grid size.
__host__ __device__ inline int col_run_base(int Zs, int rw)
{
    return Zs >= 0 ? (Zs / rw) * rw : -(((-Zs) + rw - 1) / rw) * rw;
}
__host__ __device__ inline int col_runs(int Zs, int Ze, int rw)
{
    return (Ze - col_run_base(Zs, rw) + rw - 1) / rw;
}

template <class T, int RW, int MODE>
__global__ __launch_bounds__(64 * kColBpWaves) void k_bp_col(const AgentDev *__restrict__ ags, float lambda)
{
    constexpr bool BPT = MODE == 1, NORM = MODE == 2;
    constexpr int NT = 64 * kColBpWaves, kPer = kColElems / NT;
    constexpr int kStage = 2 * kColElems, kEpi = kColBpWaves * 64 * 17 / 2;
    __shared__ float2 lds_raw[kStage > kEpi ? kStage : kEpi];
    const AgentDev &ag = ags[blockIdx.y];
    const bool p1y = ag.col_p1y != 0;
    const int pitch = ag.col_pitch;
    const int n1 = p1y ? ag.Y : ag.X, n2 = p1y ? ag.X : ag.Y;
    const int s1 = p1y ? ag.Ys : ag.Xs, s2 = p1y ? ag.Xs : ag.Ys;
    // runs start at absolute multiples of RW, so that a voxel's run (and
    // its arithmetic) does not depend on the box, e.g. on a Z-slab split
    const int zal = col_run_base(ag.Zs, RW);
    const int nt1 = (n1 + 63) / 64, nt2 = (n2 + kColBpWaves - 1) / kColBpWaves, ntz = col_runs(ag.Zs, ag.Ze, RW);
    const int nb = nt1 * nt2 * ntz, per = (nb + 7) / 8;
    const int b = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (blockIdx.x >= (unsigned)(per * 8) || b >= nb) return;
    const int bz = b % ntz, b1 = (b / ntz) % nt1, b2 = b / (ntz * nt1);
    const int lane = threadIdx.x & 63, wid = wave_id();
    const int P10 = s1 + b1 * 64, P20 = s2 + b2 * kColBpWaves, Z0 = zal + bz * RW;
    // u is stepped in FP32 from the last multiple of 32 planes at or below
    // Z0, where it is formed by the reference's FP64 chain: the same
    // arithmetic for runs of 16 and 32
    const int Z32 = Z0 >= 0 ? (Z0 / 32) * 32 : -(((-Z0) + 31) / 32) * 32;
    const int zlo = max(ag.Zs - Z0, 0);
    const int p1 = P10 + lane, p2 = P20 + wid;
    const int e1 = min(s1 + n1 - 1, P10 + 63), e2 = min(s2 + n2 - 1, P20 + kColBpWaves - 1);
    const int ew = min(RW, ag.Ze - Z0) - 1;
    const int nu = ag.nu, nv = ag.nv;
    const double y0 = ag.coef[0].c[4], yx = ag.coef[0].c[5], yy = ag.coef[0].c[6];
    // the tile's rows: v is affine in (X, Y), extremes at the corners
    double vmin = 1e300, vmax = -1e300;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c1 = (q & 1) ? e1 : P10, c2 = (q & 2) ? e2 : P20;
        const double vq = fma(yy, (double)(p1y ? c1 : c2), fma(yx, (double)(p1y ? c2 : c1), y0)) - 1.0;
        vmin = fmin(vmin, vq);
        vmax = fmax(vmax, vq);
    }
    const int klo = (int)floor(vmin), nrows = (int)floor(vmax) - klo + 2;
    // this thread's column (clamped into the box; stores are masked)
    const int q1 = min(p1, e1), q2 = min(p2, e2);
    const int X = p1y ? q2 : q1, Y = p1y ? q1 : q2;
    const double v = fma(yy, (double)Y, fma(yx, (double)X, y0)) - 1.0;
    const double vf = floor(v);
    const int k0 = (int)vf;
    const float wk0 = (unsigned)k0 < (unsigned)nv ? (float)(1.0 - (v - vf)) : 0.0f;
    const float wk1 = (unsigned)(k0 + 1) < (unsigned)nv ? (float)(v - vf) : 0.0f;
    const int lr = (k0 - klo) * pitch;
    const auto src = gview(reinterpret_cast<const T *>(BPT ? ag.p : ag.r));
    const int64_t plane = (int64_t)nu * nv;

    // S: gathered sums; D (MODE 2): summed u weight of taps off the detector
    float S[NORM ? 1 : RW], D[NORM ? RW : 1];
#pragma unroll
    for (int t = 0; t < RW; t++) {
        if constexpr (!NORM) S[t] = 0.0f;
        if constexpr (NORM) D[t] = 0.0f;
    }
    float2 pre[kPer];

    // detector columns of the tile in view a: u at the tile origin plus
    // the host's extreme offsets over the tile, one column of margin
    const int Xt = p1y ? P20 : P10, Yt = p1y ? P10 : P20;
    // (view constants come through the scalar cache, so that waiting for
    // them never waits for the staging loads in flight)
    auto window = [&](int a, const double (&uc)[4], int &fi0, int &fi1) {
        const BpSepView bv = load_bsv(ag.bsv, a);
        const double ut = fma(uc[3], (double)Z0, fma(uc[2], (double)Yt, fma(uc[1], (double)Xt, uc[0]))) - 1.0;
        fi0 = (int)floor(ut + bv.umin) - 1;
        fi1 = (int)floor(ut + bv.umax) + 1;
    };
    // rows the staging slots cover; a footprint whose slots all lie inside
    // the detector is staged without clamps or masks
    const int slot_rows = kColElems / pitch;
    const bool rows_inside = klo >= 0 && klo + slot_rows <= nv - 1;
    auto fetch = [&](int a, int fi0, int fi1) {
        const int width = fi1 - fi0 + 1;
        const uint32_t vbase = (uint32_t)(plane * a);
        if (rows_inside && fi0 >= 0 && fi0 + pitch <= nu - 1) {
            const uint32_t base = vbase + (uint32_t)klo * (uint32_t)nu + (uint32_t)fi0;
#pragma unroll
            for (int q = 0; q < kPer; q++) {
                const int e = threadIdx.x + NT * q;
                const int rr = e / pitch, cc = e - rr * pitch;
                const uint32_t o = base + (uint32_t)(rr * nu + cc);
                const float v0 = ldg<T>(src, o), v1 = ldg<T>(src, o + 1);
                pre[q] = make_float2(v0, v1 - v0);
            }
            return;
        }
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const int e = threadIdx.x + NT * q;
            const int rr = e / pitch, cc = e - rr * pitch;
            const int gk = klo + rr, gi = fi0 + cc;
            const bool ok = rr < nrows && cc < width && (unsigned)gk < (unsigned)nv;
            const bool ok0 = ok && (unsigned)gi < (unsigned)nu, ok1 = ok && (unsigned)(gi + 1) < (unsigned)nu;
            const uint32_t o = vbase + (uint32_t)min(max(gk, 0), nv - 1) * (uint32_t)nu;
            const float v0 = ldg<T>(src, o + (uint32_t)min(max(gi, 0), nu - 1));
            const float v1 = ldg<T>(src, o + (uint32_t)min(max(gi + 1, 0), nu - 1));
            const float r0 = ok0 ? v0 : 0.0f, r1 = ok1 ? v1 : 0.0f;
            pre[q] = make_float2(r0, r1 - r0);
        }
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int q = 0; q < kPer; q++) lds_raw[buf * kColElems + threadIdx.x + NT * q] = pre[q];
    };

    int fi0, fi1;
    double ucn[4];
    load_ucoef(ag.coef, 0, ucn);
    window(0, ucn, fi0, fi1);
    if constexpr (!NORM) {
        fetch(0, fi0, fi1);
        commit(0);
        __syncthreads();
    }
    for (int a = 0; a < ag.A; a++) {
        const int cfi0 = fi0, cfi1 = fi1;
        const double uc[4] = {ucn[0], ucn[1], ucn[2], ucn[3]};
        if (a + 1 < ag.A) {
            load_ucoef(ag.coef, a + 1, ucn);
            window(a + 1, ucn, fi0, fi1);
            if constexpr (!NORM) fetch(a + 1, fi0, fi1);
        }
        const double ub = fma(uc[3], (double)Z32, fma(uc[2], (double)Y, fma(uc[1], (double)X, uc[0]))) - 1.0;
        const double fb = floor(ub);
        const float f0 = (float)(ub - fb), step = (float)uc[3];
        const int ib = (int)fb - cfi0;
        const int lmax = pitch - 1;
        const float2 *row0 = lds_raw + (a & 1) * kColElems + lr + ib;
        const float2 *row1 = row0 + pitch;
        const float glo = (float)(-ib), ghi = (float)(lmax - ib) + 0.5f;
        // tap t of the run is plane Z0 + t = Z32 + OFF + t (OFF = 0, or 16
        // for the upper half of a 32-plane group with runs of 16)
        auto taps = [&](auto off_c) {
            constexpr int OFF = decltype(off_c)::value;
            if constexpr (!NORM) {
                // zero pairs outside the detector make the same update
                // exact at its side edges
#pragma unroll
                for (int t = 0; t < RW; t++) {
                    const float g = fmaf((float)(OFF + t), step, f0);
                    float du;
                    const int li = tap_clamped(g, glo, ghi, du);
                    const float2 qa = row0[li], qb = row1[li];
                    S[t] = fmaf(wk0, fmaf(du, qa.y, qa.x), S[t]);
                    S[t] = fmaf(wk1, fmaf(du, qb.y, qb.x), S[t]);
                    // at most 8 taps' LDS reads in flight (registers)
                    if ((t & 7) == 7) __builtin_amdgcn_sched_barrier(0);
                }
            } else if (!(cfi0 >= 0 && cfi1 + 1 <= nu - 1)) {
#pragma unroll
                for (int t = 0; t < RW; t++) {
                    const float g = fmaf((float)(OFF + t), step, f0);
                    const float fg = floorf(g);
                    const float du = g - fg;
                    const int i0 = ib + cfi0 + (int)fg;
                    // weight of the taps that fall off the detector's sides
                    const bool v0 = (unsigned)i0 < (unsigned)nu, v1 = (unsigned)(i0 + 1) < (unsigned)nu;
                    D[t] += v0 ? (v1 ? 0.0f : du) : (v1 ? 1.0f - du : 1.0f);
                }
            }
        };
        if (RW == 16 && Z0 != Z32) taps(std::integral_constant<int, 16>{});
        else taps(std::integral_constant<int, 0>{});
        if constexpr (!NORM) {
            if (a + 1 < ag.A) commit((a + 1) & 1);
            __syncthreads();
        }
    }
    if constexpr (NORM) __syncthreads();  // the epilogue reuses the staging LDS
    // Epilogue: the runs are stored transposed through LDS (16 lanes per 16
    // consecutive voxels of one run) as in the separable back-projection;
    // MODE 2 stores C, the others S/C of each voxel with C read back.
    float *ep = reinterpret_cast<float *>(lds_raw) + wid * (64 * 17);
    if constexpr (NORM) {
        const float kw = wk0 + wk1;
#pragma unroll
        for (int t = 0; t < RW; t++) D[t] = kw * ((float)ag.A - D[t]);
    }
    T *x = reinterpret_cast<T *>(ag.x);
    const int64_t sP1 = p1y ? ag.sy : ag.sx;
    const int X0 = p1y ? p2 : P10, Y0 = p1y ? P10 : p2;
    const int64_t jb = (int64_t)(X0 - ag.Xs) * ag.sx + (int64_t)(Y0 - ag.Ys) * ag.sy + (int64_t)(Z0 - ag.Zs) * ag.sz;
    const bool wave_ok = p2 <= e2;
    constexpr int CH = RW < 16 ? RW : 16, LPI = 64 / CH;
    static_assert(RW % CH == 0 && 64 % CH == 0, "run length must divide into chunks");
#pragma unroll
    for (int h = 0; h < RW / CH; h++) {
#pragma unroll
        for (int t = 0; t < CH; t++) {
            if constexpr (NORM) ep[lane * 17 + t] = D[CH * h + t];
            else ep[lane * 17 + t] = S[CH * h + t];
        }
        wave_lds_sync();
        const int tt = lane % CH, tw = CH * h + tt;
#pragma unroll 4
        for (int i = 0; i < CH; i++) {
            const int line = i * LPI + lane / CH;
            const float qv = ep[line * 17 + tt];
            if (wave_ok && tw >= zlo && tw <= ew && P10 + line <= e1) {
                const int64_t j = jb + line * sP1 + tw * ag.sz;
                if constexpr (NORM) {
                    ag.bpc[j] = qv;
                } else {
                    const float C = ag.bpc[j];
                    const float q = C != 0.0f ? qv / C : 0.0f;
                    if (BPT) stf<T>(x, j, q);
                    else if (C != 0.0f) stf<T>(x, j, fmaf(q, lambda, ldf<T>(x, j)));
                }
            }
        }
        wave_lds_sync();
    }
}

// ------------------------------------------------------------------------
// Column-plan back-projection on strip tiles.  A block takes one group of
// up to 64 consecutive columns of a strip (lanes) and eight runs of RW
// planes along Z (waves).  Every column of a strip lands on the same two
// detector rows k0, k0 + 1 in every view, so per view the block stages one
// (r[i], r[i+1] - r[i]) pair of each row per detector column of the
// group's footprint (the u range of the group's bounding box over the
// block's planes, at most kCsMaxW columns, checked on the host) as one
// float4, double buffered; a lane's taps read consecutive or equal
// addresses.  The arithmetic per voxel and view is k_bp_col's (u at the
// enclosing multiple of 32 planes in FP64, FP32 steps, the same tap order
// and update), so the two kernels give identical sums.  Groups with
// k0 = -2 hold the columns off the detector's rows: no view reaches them,
// their C is 0 (x = 0 at the BPT, unchanged by SIRT updates).
constexpr int kCsWaves = 8;
constexpr int kCsMaxW = 512;

int cs_max_width() { return kCsMaxW - 2; }

// Footprint of a block in one view: the detector columns over the group's
// box and the block's planes Zb0 .. Zb1, one column of margin on each side,
// from the reference's FP64 coordinate (the window only bounds what is
// staged: every tap of the block falls inside it).
static void cs_window(const double *uc, const ColGroup &g, int Zb0, int Zb1, int &fi0, int &fi1)
{
    const double ux = std::fmin(uc[1] * (double)g.xmin, uc[1] * (double)g.xmax);
    const double uy = std::fmin(uc[2] * (double)g.ymin, uc[2] * (double)g.ymax);
    const double uz = std::fmin(uc[3] * (double)Zb0, uc[3] * (double)Zb1);
    const double vx = std::fmax(uc[1] * (double)g.xmin, uc[1] * (double)g.xmax);
    const double vy = std::fmax(uc[2] * (double)g.ymin, uc[2] * (double)g.ymax);
    const double vz = std::fmax(uc[3] * (double)Zb0, uc[3] * (double)Zb1);
    fi0 = (int)std::floor(uc[0] + ux + uy + uz - 1.0) - 1;
    fi1 = (int)std::floor(uc[0] + vx + vy + vz - 1.0) + 1;
}

std::vector<int2> bp_cs_windows(const ColGroup *grp, int ngrp, const double *xc, int A, int Zs, int Ze)
{
    const int zal = col_run_base(Zs, 16), ntz = col_runs(Zs, Ze, 16), nzg = (ntz + kCsWaves - 1) / kCsWaves;
    std::vector<int2> w((size_t)ngrp * nzg * A);
    for (int gi = 0; gi < ngrp; gi++)
        for (int zg = 0; zg < nzg; zg++) {
            const int Zb0 = zal + zg * kCsWaves * 16, Zb1 = Zb0 + kCsWaves * 16 - 1;
            for (int a = 0; a < A; a++) {
                int f0, f1;
                cs_window(xc + 4 * a, grp[gi], Zb0, Zb1, f0, f1);
                w[((size_t)gi * nzg + zg) * A + a] = make_int2(f0, f1);
            }
        }
    return w;
}

template <class T, int RW, int MODE>
__global__ __launch_bounds__(64 * kCsWaves) void k_bp_cs(const AgentDev *__restrict__ ags, float lambda)
{
    constexpr bool BPT = MODE == 1, NORM = MODE == 2;
    static_assert(RW == 16, "the host's window table (bp_cs_windows) is for runs of 16 planes");
    constexpr int kEpiF4 = kCsWaves * 64 * 17 / 4;
    constexpr int kLds = 2 * kCsMaxW > kEpiF4 ? 2 * kCsMaxW : kEpiF4;
    __shared__ float4 lds4[kLds];
    __shared__ int64_t cbase[64];
    const AgentDev &ag = ags[blockIdx.y];
    const int ntz = col_runs(ag.Zs, ag.Ze, RW), nzg = (ntz + kCsWaves - 1) / kCsWaves;
    const int nb = ag.ngrp * nzg, per = (nb + 7) / 8;
    // XCD-aware order: consecutive groups (which share the strip's rows)
    // run on the same XCD's L2
    const int b = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (blockIdx.x >= (unsigned)(per * 8) || b >= nb) return;
    const int gi = b / nzg, zg = b - gi * nzg;
    const __attribute__((address_space(4))) int *gp =
        (const __attribute__((address_space(4))) int *)(ag.col_grp + gi);
    const int gj0 = gp[0], gn = gp[1], k0 = gp[6];
    const int lane = threadIdx.x & 63, wid = wave_id();
    const int zal = col_run_base(ag.Zs, RW);
    const int bz = zg * kCsWaves + wid;
    const int Z0 = zal + bz * RW;
    const int Z32 = Z0 >= 0 ? (Z0 / 32) * 32 : -(((-Z0) + 31) / 32) * 32;
    const int zlo = max(ag.Zs - Z0, 0);
    const int ew = min(RW, ag.Ze - Z0) - 1;
    const bool wave_ok = bz < ntz;
    const int nu = ag.nu, nv = ag.nv;
    // this lane's column (the group's last one for lanes past its end;
    // stores are masked)
    const int j = gj0 + min(lane, gn - 1);
    const int2 xy = ag.col_xy[j];
    const float2 cw = ag.col_w[j];
    const int X = xy.x, Y = xy.y;
    const float wk0 = (unsigned)k0 < (unsigned)nv ? cw.x : 0.0f;
    const float wk1 = (unsigned)(k0 + 1) < (unsigned)nv ? cw.y : 0.0f;
    if (wid == 0) cbase[lane] = (int64_t)(X - ag.Xs) * ag.sx + (int64_t)(Y - ag.Ys) * ag.sy;
    const auto src = gview(reinterpret_cast<const T *>(BPT ? ag.p : ag.r));
    const int64_t plane = (int64_t)nu * nv;
    const bool row0_ok = (unsigned)k0 < (unsigned)nv, row1_ok = (unsigned)(k0 + 1) < (unsigned)nv;

    float S[NORM ? 1 : RW], D[NORM ? RW : 1];
#pragma unroll
    for (int t = 0; t < RW; t++) {
        if constexpr (!NORM) S[t] = 0.0f;
        if constexpr (NORM) D[t] = 0.0f;
    }
    // footprint of view a (the host's cs_window), through the scalar cache
    const __attribute__((address_space(4))) int *wtab =
        (const __attribute__((address_space(4))) int *)(ag.cs_win + ((int64_t)gi * nzg + zg) * ag.A);
    auto window = [&](int a, int &fi0, int &fi1) {
        fi0 = wtab[2 * a];
        fi1 = wtab[2 * a + 1];
    };
    float4 pre;
    auto fetch = [&](int a, int fi0, int fi1) {
        const int e = threadIdx.x;
        const int gi0 = fi0 + e;
        const bool in = e <= fi1 - fi0;
        const bool ok0 = in && (unsigned)gi0 < (unsigned)nu, ok1 = in && (unsigned)(gi0 + 1) < (unsigned)nu;
        const uint32_t vb = (uint32_t)(plane * a);
        const uint32_t i0 = (uint32_t)min(max(gi0, 0), nu - 1), i1 = (uint32_t)min(max(gi0 + 1, 0), nu - 1);
        const uint32_t o0 = vb + (uint32_t)min(max(k0, 0), nv - 1) * (uint32_t)nu;
        const uint32_t o1 = vb + (uint32_t)min(max(k0 + 1, 0), nv - 1) * (uint32_t)nu;
        const float a0 = ok0 && row0_ok ? ldg<T>(src, o0 + i0) : 0.0f, a1 = ok1 && row0_ok ? ldg<T>(src, o0 + i1) : 0.0f;
        const float b0 = ok0 && row1_ok ? ldg<T>(src, o1 + i0) : 0.0f, b1 = ok1 && row1_ok ? ldg<T>(src, o1 + i1) : 0.0f;
        pre = make_float4(a0, a1 - a0, b0, b1 - b0);
    };
    int fi0, fi1;
    double ucn[4];
    load_ucoef(ag.coef, 0, ucn);
    window(0, fi0, fi1);
    // a wave whose 64 staging slots all lie past the footprint's width
    // stages nothing (taps never read past it)
    bool stage = wid * 64 <= fi1 - fi0;
    if constexpr (!NORM) {
        if (stage) {
            fetch(0, fi0, fi1);
            lds4[threadIdx.x] = pre;
        }
    }
    __syncthreads();
    const int nview = k0 < -1 ? 0 : ag.A;  // block-uniform
    for (int a = 0; a < nview; a++) {
        const int cfi0 = fi0, cfi1 = fi1;
        const double uc[4] = {ucn[0], ucn[1], ucn[2], ucn[3]};
        if (a + 1 < nview) {
            load_ucoef(ag.coef, a + 1, ucn);
            window(a + 1, fi0, fi1);
            stage = wid * 64 <= fi1 - fi0;
            if constexpr (!NORM)
                if (stage) fetch(a + 1, fi0, fi1);
        }
        const double ub = fma(uc[3], (double)Z32, fma(uc[2], (double)Y, fma(uc[1], (double)X, uc[0]))) - 1.0;
        const double fb = floor(ub);
        const float f0 = (float)(ub - fb), step = (float)uc[3];
        const int ib = (int)fb - cfi0;
        const float4 *stg = lds4 + (a & 1) * kCsMaxW + ib;
        auto taps = [&](auto off_c) {
            constexpr int OFF = decltype(off_c)::value;
            if constexpr (!NORM) {
#pragma unroll
                for (int t = 0; t < RW; t++) {
                    const float g = fmaf((float)(OFF + t), step, f0);
                    // no clamp: the window covers the block's box over all its
                    // planes with a column of margin, far more than the FP32
                    // steps of g can drift from the FP64 corners
                    const float du = __builtin_amdgcn_fractf(g);
                    const int li = cvt_flr(g);
                    const float4 q = stg[li];
                    S[t] = fmaf(wk0, fmaf(du, q.y, q.x), S[t]);
                    S[t] = fmaf(wk1, fmaf(du, q.w, q.z), S[t]);
                    if ((t & 7) == 7) __builtin_amdgcn_sched_barrier(0);
                }
            } else if (!(cfi0 >= 0 && cfi1 + 1 <= nu - 1)) {
#pragma unroll
                for (int t = 0; t < RW; t++) {
                    const float g = fmaf((float)(OFF + t), step, f0);
                    const float fg = floorf(g);
                    const float du = g - fg;
                    const int i0 = ib + cfi0 + (int)fg;
                    const bool v0 = (unsigned)i0 < (unsigned)nu, v1 = (unsigned)(i0 + 1) < (unsigned)nu;
                    D[t] += v0 ? (v1 ? 0.0f : du) : (v1 ? 1.0f - du : 1.0f);
                }
            }
        };
        if (RW == 16 && Z0 != Z32) taps(std::integral_constant<int, 16>{});
        else taps(std::integral_constant<int, 0>{});
        if constexpr (!NORM) {
            if (a + 1 < nview && stage) lds4[((a + 1) & 1) * kCsMaxW + threadIdx.x] = pre;
            __syncthreads();
        }
    }
    if (NORM || nview == 0) __syncthreads();  // cbase, and the epilogue reuses the staging LDS
    float *ep = reinterpret_cast<float *>(lds4) + wid * (64 * 17);
    if constexpr (NORM) {
        const float kw = wk0 + wk1;
#pragma unroll
        for (int t = 0; t < RW; t++) D[t] = kw * ((float)ag.A - D[t]);
    }
    T *x = reinterpret_cast<T *>(ag.x);
    const int64_t zb = (int64_t)(Z0 - ag.Zs) * ag.sz;
    constexpr int CH = RW < 16 ? RW : 16, LPI = 64 / CH;
    static_assert(RW % CH == 0 && 64 % CH == 0, "run length must divide into chunks");
#pragma unroll
    for (int h = 0; h < RW / CH; h++) {
#pragma unroll
        for (int t = 0; t < CH; t++) {
            if constexpr (NORM) ep[lane * 17 + t] = D[CH * h + t];
            else ep[lane * 17 + t] = S[CH * h + t];
        }
        wave_lds_sync();
        const int tt = lane % CH, tw = CH * h + tt;
#pragma unroll 4
        for (int i = 0; i < CH; i++) {
            const int line = i * LPI + lane / CH;
            const float qv = ep[line * 17 + tt];
            if (wave_ok && tw >= zlo && tw <= ew && line < gn) {
                const int64_t jv = cbase[line] + zb + tw * ag.sz;
                if constexpr (NORM) {
                    ag.bpc[jv] = qv;
                } else {
                    const float C = ag.bpc[jv];
                    const float q = C != 0.0f ? qv / C : 0.0f;
                    if (BPT) stf<T>(x, jv, q);
                    else if (C != 0.0f) stf<T>(x, jv, fmaf(q, lambda, ldf<T>(x, jv)));
                }
            }
        }
        wave_lds_sync();
    }
}

// ------------------------------------------------------------------------
// Fused consensus step over all K agents of a solve (element-wise, HBM
// bound):  z_a = 2 wbar - w_a (the state this step started from);
// w_a = rho (2 v_a - z_a) + (1 - rho) w_a; wbar' = sum_a w_a / K;
// next working volume x_a = 2 wbar' - w_a; FP64 sum of (w_a - wbar')^2.
// Arithmetic follows the reference's host loops exactly (no contraction).
struct ConsArgs {
    int K;
    int64_t N;
    float rho;
    int write_next;
    void *w[kMaxAgents];
    void *x[kMaxAgents];
    void *wbar;
    double *partial;
};

template <class T>
__global__ __launch_bounds__(256) void k_consensus(ConsArgs ca)
{
#pragma clang fp contract(off)
    double d2 = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < ca.N; j += stride) {
        const float wb = ldf<T>(reinterpret_cast<const T *>(ca.wbar), j);
        float wn[kMaxAgents];
        float acc = 0.f;
#pragma unroll
        for (int a = 0; a < kMaxAgents; a++) {
            if (a >= ca.K) break;
            const float wo = ldf<T>(reinterpret_cast<const T *>(ca.w[a]), j);
            const float vv = ldf<T>(reinterpret_cast<const T *>(ca.x[a]), j);
            const float z = 2.0f * wb - wo;
            const float t1 = ca.rho * (2.0f * vv - z);
            const float t2 = (1.0f - ca.rho) * wo;
            // Round w_a to storage before it enters the mean, as a stored
            // state would be.
            stf<T>(reinterpret_cast<T *>(ca.w[a]), j, t1 + t2);
            wn[a] = ldf<T>(reinterpret_cast<const T *>(ca.w[a]), j);
            acc = (a == 0) ? wn[a] : acc + wn[a];
        }
        const float wbn_f = acc / (float)ca.K;
        stf<T>(reinterpret_cast<T *>(ca.wbar), j, wbn_f);
        const float wbn = ldf<T>(reinterpret_cast<const T *>(ca.wbar), j);
#pragma unroll
        for (int a = 0; a < kMaxAgents; a++) {
            if (a >= ca.K) break;
            if (ca.write_next) stf<T>(reinterpret_cast<T *>(ca.x[a]), j, 2.0f * wbn - wn[a]);
            const double e = (double)wn[a] - (double)wbn;
            d2 += e * e;
        }
    }
    __shared__ double sh[4];
    const double t = block_sum<256>(d2, sh);
    if (threadIdx.x == 0) ca.partial[blockIdx.x] = t;
}

// ------------------------------------------------------------------------
// Stand-alone element-wise kernels (single agent, the reference's
// per-call structure).
template <class T>
__global__ void k_form_z(int64_t N, const T *__restrict__ w, const T *__restrict__ wbar,
                         T *__restrict__ z)
{
#pragma clang fp contract(off)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += stride)
        stf<T>(z, j, 2.0f * ldf<T>(wbar, j) - ldf<T>(w, j));
}

template <class T>
__global__ void k_mann(int64_t N, const T *__restrict__ v, T *__restrict__ w, const T *__restrict__ z,
                       float rho)
{
#pragma clang fp contract(off)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += stride) {
        const float t1 = rho * (2.0f * ldf<T>(v, j) - ldf<T>(z, j));
        const float t2 = (1.0f - rho) * ldf<T>(w, j);
        stf<T>(w, j, t1 + t2);
    }
}

// Mann update with z recomputed from the step's starting state:
// w = rho (2v - (2 wbar - w)) + (1 - rho) w.
template <class T>
__global__ void k_mann_z(int64_t N, const T *__restrict__ v, T *__restrict__ w,
                         const T *__restrict__ wbar, float rho)
{
#pragma clang fp contract(off)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < N; j += stride) {
        const float wo = ldf<T>(w, j);
        const float z = 2.0f * ldf<T>(wbar, j) - wo;
        const float t1 = rho * (2.0f * ldf<T>(v, j) - z);
        const float t2 = (1.0f - rho) * wo;
        stf<T>(w, j, t1 + t2);
    }
}

// MACE initialisation: B = sum_a BPT_a (agent order, FP32, as the SUM
// reduction of MultiAxis.cpp:192 with the oracle's order), then
// w_a = wbar = B and the first working volumes z_a = 2B - B = B.
struct InitArgs {
    int K;
    int64_t N;
    void *w[kMaxAgents];
    void *x[kMaxAgents];
    void *wbar;
};

template <class T>
__global__ void k_init_sum(InitArgs ia)
{
#pragma clang fp contract(off)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < ia.N; j += stride) {
        float acc = 0.0f;
        for (int a = 0; a < ia.K; a++) acc += ldf<T>(reinterpret_cast<const T *>(ia.x[a]), j);
        stf<T>(reinterpret_cast<T *>(ia.wbar), j, acc);
        const float b = ldf<T>(reinterpret_cast<const T *>(ia.wbar), j);
        for (int a = 0; a < ia.K; a++) {
            stf<T>(reinterpret_cast<T *>(ia.w[a]), j, b);
            stf<T>(reinterpret_cast<T *>(ia.x[a]), j, 2.0f * b - b);
        }
    }
}

struct MeanArgs {
    int K;
    int64_t N;
    const void *w[kMaxAgents];
    void *wbar;
};

template <class T>
__global__ void k_mean(MeanArgs ma)
{
#pragma clang fp contract(off)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < ma.N; j += stride) {
        float acc = ldf<T>(reinterpret_cast<const T *>(ma.w[0]), j);
        for (int a = 1; a < ma.K; a++) acc += ldf<T>(reinterpret_cast<const T *>(ma.w[a]), j);
        stf<T>(reinterpret_cast<T *>(ma.wbar), j, acc / (float)ma.K);
    }
}

template <class T>
__global__ void k_residual(int64_t P, const T *__restrict__ p, const float *__restrict__ s,
                           const float *__restrict__ c, T *__restrict__ r)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < P; j += stride) {
        float sv = s[j];
        if (c[j] != 0.0f) sv = sv / c[j];
        stf<T>(r, j, ldf<T>(p, j) - sv);
    }
}

// Storage conversion (float -> T and T -> float).
template <class T>
__global__ void k_to_storage(int64_t n, const float *__restrict__ src, T *__restrict__ dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride)
        stf<T>(dst, j, src[j]);
}
template <class T>
__global__ void k_from_storage(int64_t n, const T *__restrict__ src, float *__restrict__ dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride)
        dst[j] = ldf<T>(src, j);
}
template <class T>
__global__ void k_fill(int64_t n, T *__restrict__ dst, float v)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride)
        stf<T>(dst, j, v);
}

// Volume layout change: element (X, Y, Z) moves from src[X*ss.x + Y*ss.y +
// Z*ss.z] to dst[X*ds.x + Y*ds.y + Z*ds.z].  A is the axis contiguous in
// src, B the axis contiguous in dst; 32 x 32 tiles of the (A, B) plane
// pass through LDS so that both the reads and the writes are coalesced.
struct PermArgs {
    int n[3];
    int64_t ss[3], ds[3];
    int A, B, C;
};

template <class T>
__global__ __launch_bounds__(256) void k_permute(const T *__restrict__ src, T *__restrict__ dst, PermArgs pa)
{
    __shared__ float tile[32][33];
    const int nA = pa.n[pa.A], nB = pa.n[pa.B], nC = pa.n[pa.C];
    const int ta = (nA + 31) / 32, tb = (nB + 31) / 32;
    int64_t b = blockIdx.x;
    if (b >= (int64_t)ta * tb * nC) return;
    const int ia = (int)(b % ta);
    b /= ta;
    const int ib = (int)(b % tb);
    const int c = (int)(b / tb);
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int a0 = ia * 32, b0 = ib * 32;
    const int64_t sA = pa.ss[pa.A], sB = pa.ss[pa.B], sC = pa.ss[pa.C];
    const int64_t dA = pa.ds[pa.A], dB = pa.ds[pa.B], dC = pa.ds[pa.C];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int bb = b0 + ty + 8 * q, aa = a0 + tx;
        if (aa < nA && bb < nB) tile[ty + 8 * q][tx] = ldf<T>(src, aa * sA + bb * sB + c * sC);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int aa = a0 + ty + 8 * q, bb = b0 + tx;
        if (aa < nA && bb < nB) stf<T>(dst, aa * dA + bb * dB + c * dC, tile[tx][ty + 8 * q]);
    }
}

// Deterministic fixed-order sum of n FP64 partials (one block of 256).
__global__ __launch_bounds__(256) void k_reduce(const double *__restrict__ part, int64_t n,
                                               double *__restrict__ out)
{
    double v = 0.0;
    for (int64_t j = threadIdx.x; j < n; j += 256) v += part[j];
    __shared__ double sh[256];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// ------------------------------------------------------------------------
// Plain form: the reference's kernel structure on this GPU.  One thread
// per voxel in 8x8x8 blocks; the forward projection splats with float
// atomics into zeroed s and c (recomputed every iteration), and the
// back-projection gathers straight from HBM accumulating S and C in FP64.
__global__ __launch_bounds__(512) void k_plain_splat(AgentDev ag, float *__restrict__ s,
                                                    float *__restrict__ c)
{
    const int X = ag.Xs + blockIdx.x * 8 + threadIdx.x;
    const int Y = ag.Ys + blockIdx.y * 8 + threadIdx.y;
    const int Z = ag.Zs + blockIdx.z * 8 + threadIdx.z;
    if (X >= ag.Xe || Y >= ag.Ye || Z >= ag.Ze) return;
    const float *x = reinterpret_cast<const float *>(ag.x);
    const double xv = (double)x[(int64_t)(X - ag.Xs) * ag.sx + (int64_t)(Y - ag.Ys) * ag.sy +
                                (int64_t)(Z - ag.Zs) * ag.sz];
    const int nu = ag.nu, nv = ag.nv;
    for (int a = 0; a < ag.A; a++) {
        const ViewCoef vc = ag.coef[a];
        const double u = coord(vc.c, X, Y, Z), v = coord(vc.c + 4, X, Y, Z);
        const double fu = floor(u), fv = floor(v);
        const int i0 = (int)fu, k0 = (int)fv;
        const double du = u - fu, dv = v - fv;
        const int64_t o = (int64_t)a * nu * nv;
        const double w[4] = {(1 - du) * (1 - dv), du * (1 - dv), (1 - du) * dv, du * dv};
        const int ii[4] = {i0, i0 + 1, i0, i0 + 1};
        const int kk[4] = {k0, k0, k0 + 1, k0 + 1};
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if ((unsigned)ii[t] < (unsigned)nu && (unsigned)kk[t] < (unsigned)nv) {
                const int64_t n = o + (int64_t)kk[t] * nu + ii[t];
                atomicAdd(&s[n], (float)(w[t] * xv));
                atomicAdd(&c[n], (float)w[t]);
            }
        }
    }
}

__global__ __launch_bounds__(512) void k_plain_bp(AgentDev ag, const float *__restrict__ r,
                                                 float lambda, int bpt)
{
    const int X = ag.Xs + blockIdx.x * 8 + threadIdx.x;
    const int Y = ag.Ys + blockIdx.y * 8 + threadIdx.y;
    const int Z = ag.Zs + blockIdx.z * 8 + threadIdx.z;
    if (X >= ag.Xe || Y >= ag.Ye || Z >= ag.Ze) return;
    const int nu = ag.nu, nv = ag.nv;
    double S = 0.0, C = 0.0;
    for (int a = 0; a < ag.A; a++) {
        const ViewCoef vc = ag.coef[a];
        const double u = coord(vc.c, X, Y, Z), v = coord(vc.c + 4, X, Y, Z);
        const double fu = floor(u), fv = floor(v);
        const int i0 = (int)fu, k0 = (int)fv;
        const double du = u - fu, dv = v - fv;
        const float *rv = r + (int64_t)a * nu * nv;
        const double w[4] = {(1 - du) * (1 - dv), du * (1 - dv), (1 - du) * dv, du * dv};
        const int ii[4] = {i0, i0 + 1, i0, i0 + 1};
        const int kk[4] = {k0, k0, k0 + 1, k0 + 1};
#pragma unroll
        for (int t = 0; t < 4; t++) {
            if ((unsigned)ii[t] < (unsigned)nu && (unsigned)kk[t] < (unsigned)nv) {
                S = fma(w[t], (double)rv[(int64_t)kk[t] * nu + ii[t]], S);
                C += w[t];
            }
        }
    }
    float *x = reinterpret_cast<float *>(ag.x);
    const int64_t j = (int64_t)(X - ag.Xs) * ag.sx + (int64_t)(Y - ag.Ys) * ag.sy + (int64_t)(Z - ag.Zs) * ag.sz;
    if (bpt) {
        x[j] = C != 0.0 ? (float)(S / C) : 0.0f;
    } else if (C != 0.0) {
        x[j] = (float)fma((double)(float)(S / C), (double)lambda, (double)x[j]);
    }
}

// ========================================================================
// launchers
namespace {
inline int ew_grid(int64_t n)
{
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (int)g;
}
}  // namespace

int fp_blocks(const AgentDev &ag, bool lanes_k)
{
    const int TI = lanes_k ? 4 : 64, TK = lanes_k ? 64 : 4;
    return ((ag.nu + TI - 1) / TI) * ((ag.nv + TK - 1) / TK) * ag.A;
}

int bp_blocks(const AgentDev &ag, int tz)
{
    return ((ag.X + 63) / 64) * ((ag.Y + 3) / 4) * ((ag.Z + tz - 1) / tz);
}

template <class T>
static hipError_t launch_fp_t(const AgentDev *d_ags, int nag, int max_blocks, int mode, bool lanes_k,
                              float *s_out, float *c_out, hipStream_t st)
{
    dim3 grid(max_blocks, nag), block(256);
#define MTR_FP(M, LK) hipLaunchKernelGGL((k_fp<T, M, LK>), grid, block, 0, st, d_ags, s_out, c_out)
    if (lanes_k) {
        if (mode == 0) MTR_FP(0, true);
        else if (mode == 1) MTR_FP(1, true);
        else MTR_FP(2, true);
    } else {
        if (mode == 0) MTR_FP(0, false);
        else if (mode == 1) MTR_FP(1, false);
        else MTR_FP(2, false);
    }
#undef MTR_FP
    return hipGetLastError();
}

hipError_t launch_fp(int prec, const AgentDev *d_ags, int nag, int max_blocks, int mode, bool lanes_k,
                     float *s_out, float *c_out, hipStream_t st)
{
    if (prec == MTR_PREC_BF16)
        return launch_fp_t<__hip_bfloat16>(d_ags, nag, max_blocks, mode, lanes_k, s_out, c_out, st);
    if (prec == MTR_PREC_FP16)
        return launch_fp_t<__half>(d_ags, nag, max_blocks, mode, lanes_k, s_out, c_out, st);
    return launch_fp_t<float>(d_ags, nag, max_blocks, mode, lanes_k, s_out, c_out, st);
}

template <class T>
static hipError_t launch_bp_t(const AgentDev *d_ags, int nag, int max_blocks, int tz, bool bpt,
                              float lambda, const void *src, int lds_cap, hipStream_t st)
{
    dim3 grid(max_blocks, nag), block(256);
    const size_t shm = (size_t)lds_cap * sizeof(float);
    const T *s = reinterpret_cast<const T *>(src);
#define MTR_BP(TZ, B) hipLaunchKernelGGL((k_bp<T, TZ, B>), grid, block, shm, st, d_ags, lambda, s, lds_cap)
    if (tz == 4) {
        if (bpt) MTR_BP(4, true);
        else MTR_BP(4, false);
    } else if (tz == 16) {
        if (bpt) MTR_BP(16, true);
        else MTR_BP(16, false);
    } else {
        if (bpt) MTR_BP(8, true);
        else MTR_BP(8, false);
    }
#undef MTR_BP
    return hipGetLastError();
}

hipError_t launch_bp(int prec, const AgentDev *d_ags, int nag, int max_blocks, int tz, bool bpt,
                     float lambda, const void *src, int lds_cap, hipStream_t st)
{
    if (prec == MTR_PREC_BF16)
        return launch_bp_t<__hip_bfloat16>(d_ags, nag, max_blocks, tz, bpt, lambda, src, lds_cap, st);
    if (prec == MTR_PREC_FP16)
        return launch_bp_t<__half>(d_ags, nag, max_blocks, tz, bpt, lambda, src, lds_cap, st);
    return launch_bp_t<float>(d_ags, nag, max_blocks, tz, bpt, lambda, src, lds_cap, st);
}

int fp_stream_blocks(const AgentDev &ag, int g)
{
    return (((ag.nu + 4 * g - 1) / (4 * g)) * ((ag.nv + 63) / 64) * ag.nseg * ag.A + 7) / 8 * 8;
}

template <class T>
static hipError_t launch_fp_seg_res_t(const AgentDev *d_ags, int nag, int blocks, hipStream_t st)
{
    dim3 grid(blocks, nag), block(256);
    hipLaunchKernelGGL((k_fp_seg_res<T>), grid, block, 0, st, d_ags);
    return hipGetLastError();
}

hipError_t launch_fp_seg_res(int prec, const AgentDev *d_ags, int nag, int blocks, hipStream_t st)
{
    if (prec == MTR_PREC_BF16) return launch_fp_seg_res_t<__hip_bfloat16>(d_ags, nag, blocks, st);
    if (prec == MTR_PREC_FP16) return launch_fp_seg_res_t<__half>(d_ags, nag, blocks, st);
    return launch_fp_seg_res_t<float>(d_ags, nag, blocks, st);
}

template <class T, int G>
static hipError_t launch_fp_stream_g(const AgentDev *d_ags, int nag, int max_blocks, int nc, hipStream_t st)
{
    dim3 grid(max_blocks, nag), block(256);
    if (nc == 2) hipLaunchKernelGGL((k_fp_stream<T, 2, G>), grid, block, 0, st, d_ags);
    else if (nc == 3) hipLaunchKernelGGL((k_fp_stream<T, 3, G>), grid, block, 0, st, d_ags);
    else hipLaunchKernelGGL((k_fp_stream<T, 4, G>), grid, block, 0, st, d_ags);
    return hipGetLastError();
}

template <class T>
static hipError_t launch_fp_stream_t(const AgentDev *d_ags, int nag, int max_blocks, int nc, int g, hipStream_t st)
{
    if (g == 2) return launch_fp_stream_g<T, 2>(d_ags, nag, max_blocks, nc, st);
    return launch_fp_stream_g<T, 1>(d_ags, nag, max_blocks, nc, st);
}

hipError_t launch_fp_stream(int prec, const AgentDev *d_ags, int nag, int max_blocks, int nc, int g, hipStream_t st)
{
    if (prec == MTR_PREC_BF16) return launch_fp_stream_t<__hip_bfloat16>(d_ags, nag, max_blocks, nc, g, st);
    if (prec == MTR_PREC_FP16) return launch_fp_stream_t<__half>(d_ags, nag, max_blocks, nc, g, st);
    return launch_fp_stream_t<float>(d_ags, nag, max_blocks, nc, g, st);
}

// odd (in float2 units) so that a column of 32 rows written by 32 lanes
// spreads over all LDS banks; 79 so that a footprint of 32 rows stages in
// five slots per thread (32 x 79 <= 5 x 512), which with the register cap
// below fits three blocks per CU without spills
constexpr int kSepPitch = 79;
// views staged per barrier: two for runs of up to 16 planes (the four
// footprints then still fit three blocks per CU in LDS), one for 32
template <int RW>
constexpr int kSepVps = RW <= 16 ? 2 : 1;

int bp_sep_blocks(const AgentDev &ag, int W, int P1, int P2, int rw)
{
    auto ext = [&](int ax) { return ax == 0 ? ag.X : (ax == 1 ? ag.Y : ag.Z); };
    const int nb = ((ext(P1) + 63) / 64) * ((ext(P2) + kSepWaves - 1) / kSepWaves) * ((ext(W) + rw - 1) / rw);
    return ((nb + 7) / 8) * 8;
}

int bp_sep_pitch() { return kSepPitch; }
int bp_sep_waves() { return kSepWaves; }

// Footprint of a separable BP tile (64 x kSepWaves lines from P10, P20) in
// one view: detector columns from the reference's FP64 coordinate at the
// tile's corner and the view's u range over the tile, with the margins the
// staging needs, clamped to [-1, nu].
static void sep_window(const BpSepView &bv, int P10, int P20, int nu, int &fi0, int &fi1)
{
    const double u0 = std::fma(bv.a1, (double)P10, std::fma(bv.a2, (double)P20, bv.c0));
    fi0 = std::max((int)std::floor(u0 + bv.umin) - 1, -1);
    fi1 = std::min((int)std::floor(u0 + bv.umax) + 2, nu);
}

std::vector<int2> bp_sep_windows(const BpSepView *bsv, int A, const AgentDev &ag, int P1, int P2)
{
    auto ext = [&](int ax) { return ax == 0 ? ag.X : (ax == 1 ? ag.Y : ag.Z); };
    auto start = [&](int ax) { return ax == 0 ? ag.Xs : (ax == 1 ? ag.Ys : ag.Zs); };
    const int nt1 = (ext(P1) + 63) / 64, nt2 = (ext(P2) + kSepWaves - 1) / kSepWaves;
    std::vector<int2> w((size_t)nt1 * nt2 * A);
    for (int b2 = 0; b2 < nt2; b2++)
        for (int b1 = 0; b1 < nt1; b1++)
            for (int a = 0; a < A; a++) {
                int f0, f1;
                sep_window(bsv[a], start(P1) + b1 * 64, start(P2) + b2 * kSepWaves, ag.nu, f0, f1);
                w[((size_t)b1 + (size_t)nt1 * b2) * A + a] = make_int2(f0, f1);
            }
    return w;
}

template <class T, int RW>
static void launch_bp_sep_rw(dim3 grid, const AgentDev *d_ags, bool bpt, float lambda, BpSepArgs sa, hipStream_t st)
{
    if (bpt)
        hipLaunchKernelGGL((k_bp_sep<T, RW, kSepPitch, true, kSepVps<RW>>), grid, dim3(64 * kSepWaves), 0, st, d_ags,
                           lambda, sa);
    else
        hipLaunchKernelGGL((k_bp_sep<T, RW, kSepPitch, false, kSepVps<RW>>), grid, dim3(64 * kSepWaves), 0, st, d_ags,
                           lambda, sa);
}

template <class T>
static hipError_t launch_bp_sep_t(const AgentDev *d_ags, int nag, int max_blocks, bool bpt, float lambda, int W,
                                  int P1, int P2, int rw, hipStream_t st)
{
    BpSepArgs sa{W, P1, P2};
    dim3 grid(max_blocks, nag);
    if (rw == 32) launch_bp_sep_rw<T, 32>(grid, d_ags, bpt, lambda, sa, st);
    else if (rw == 8) launch_bp_sep_rw<T, 8>(grid, d_ags, bpt, lambda, sa, st);
    else launch_bp_sep_rw<T, 16>(grid, d_ags, bpt, lambda, sa, st);
    return hipGetLastError();
}

hipError_t launch_bp_sep(int prec, const AgentDev *d_ags, int nag, int max_blocks, bool bpt, float lambda, int W,
                         int P1, int P2, int rw, hipStream_t st)
{
    if (prec == MTR_PREC_BF16)
        return launch_bp_sep_t<__hip_bfloat16>(d_ags, nag, max_blocks, bpt, lambda, W, P1, P2, rw, st);
    if (prec == MTR_PREC_FP16) return launch_bp_sep_t<__half>(d_ags, nag, max_blocks, bpt, lambda, W, P1, P2, rw, st);
    return launch_bp_sep_t<float>(d_ags, nag, max_blocks, bpt, lambda, W, P1, P2, rw, st);
}

int consensus_blocks(int64_t N) { return ew_grid(N) > 8192 ? 8192 : ew_grid(N); }

hipError_t launch_consensus(int prec, int K, int64_t N, float rho, int write_next, void *const *w,
                            void *const *x, void *wbar, double *partial, int blocks, hipStream_t st)
{
    ConsArgs ca;
    ca.K = K;
    ca.N = N;
    ca.rho = rho;
    ca.write_next = write_next;
    for (int a = 0; a < kMaxAgents; a++) {
        ca.w[a] = a < K ? w[a] : nullptr;
        ca.x[a] = a < K ? x[a] : nullptr;
    }
    ca.wbar = wbar;
    ca.partial = partial;
    if (prec == MTR_PREC_BF16)
        hipLaunchKernelGGL(k_consensus<__hip_bfloat16>, dim3(blocks), dim3(256), 0, st, ca);
    else if (prec == MTR_PREC_FP16)
        hipLaunchKernelGGL(k_consensus<__half>, dim3(blocks), dim3(256), 0, st, ca);
    else
        hipLaunchKernelGGL(k_consensus<float>, dim3(blocks), dim3(256), 0, st, ca);
    return hipGetLastError();
}

#define MTR_DISPATCH(prec, KERNEL, ...)                                                     \
    do {                                                                                    \
        if ((prec) == MTR_PREC_BF16)                                                        \
            hipLaunchKernelGGL(KERNEL<__hip_bfloat16>, __VA_ARGS__);                        \
        else if ((prec) == MTR_PREC_FP16)                                                   \
            hipLaunchKernelGGL(KERNEL<__half>, __VA_ARGS__);                                \
        else                                                                                \
            hipLaunchKernelGGL(KERNEL<float>, __VA_ARGS__);                                 \
    } while (0)

hipError_t launch_form_z(int prec, int64_t N, const void *w, const void *wbar, void *z, hipStream_t st)
{
    if (prec == MTR_PREC_BF16)
        hipLaunchKernelGGL(k_form_z<__hip_bfloat16>, dim3(ew_grid(N)), dim3(256), 0, st, N,
                           (const __hip_bfloat16 *)w, (const __hip_bfloat16 *)wbar, (__hip_bfloat16 *)z);
    else if (prec == MTR_PREC_FP16)
        hipLaunchKernelGGL(k_form_z<__half>, dim3(ew_grid(N)), dim3(256), 0, st, N, (const __half *)w,
                           (const __half *)wbar, (__half *)z);
    else
        hipLaunchKernelGGL(k_form_z<float>, dim3(ew_grid(N)), dim3(256), 0, st, N, (const float *)w,
                           (const float *)wbar, (float *)z);
    return hipGetLastError();
}

hipError_t launch_mann(int prec, int64_t N, const void *v, void *w, const void *z, float rho,
                       hipStream_t st)
{
    if (prec == MTR_PREC_BF16)
        hipLaunchKernelGGL(k_mann<__hip_bfloat16>, dim3(ew_grid(N)), dim3(256), 0, st, N,
                           (const __hip_bfloat16 *)v, (__hip_bfloat16 *)w, (const __hip_bfloat16 *)z, rho);
    else if (prec == MTR_PREC_FP16)
        hipLaunchKernelGGL(k_mann<__half>, dim3(ew_grid(N)), dim3(256), 0, st, N, (const __half *)v,
                           (__half *)w, (const __half *)z, rho);
    else
        hipLaunchKernelGGL(k_mann<float>, dim3(ew_grid(N)), dim3(256), 0, st, N, (const float *)v,
                           (float *)w, (const float *)z, rho);
    return hipGetLastError();
}

hipError_t launch_mann_z(int prec, int64_t N, const void *v, void *w, const void *wbar, float rho,
                         hipStream_t st)
{
    if (prec == MTR_PREC_BF16)
        hipLaunchKernelGGL(k_mann_z<__hip_bfloat16>, dim3(ew_grid(N)), dim3(256), 0, st, N,
                           (const __hip_bfloat16 *)v, (__hip_bfloat16 *)w, (const __hip_bfloat16 *)wbar, rho);
    else if (prec == MTR_PREC_FP16)
        hipLaunchKernelGGL(k_mann_z<__half>, dim3(ew_grid(N)), dim3(256), 0, st, N, (const __half *)v,
                           (__half *)w, (const __half *)wbar, rho);
    else
        hipLaunchKernelGGL(k_mann_z<float>, dim3(ew_grid(N)), dim3(256), 0, st, N, (const float *)v,
                           (float *)w, (const float *)wbar, rho);
    return hipGetLastError();
}

hipError_t launch_init_sum(int prec, int K, int64_t N, void *const *w, void *const *x, void *wbar,
                           hipStream_t st)
{
    InitArgs ia;
    ia.K = K;
    ia.N = N;
    for (int a = 0; a < kMaxAgents; a++) {
        ia.w[a] = a < K ? w[a] : nullptr;
        ia.x[a] = a < K ? x[a] : nullptr;
    }
    ia.wbar = wbar;
    MTR_DISPATCH(prec, k_init_sum, dim3(ew_grid(N)), dim3(256), 0, st, ia);
    return hipGetLastError();
}

hipError_t launch_mean(int prec, int K, int64_t N, const void *const *w, void *wbar, hipStream_t st)
{
    MeanArgs ma;
    ma.K = K;
    ma.N = N;
    for (int a = 0; a < kMaxAgents; a++) ma.w[a] = a < K ? w[a] : nullptr;
    ma.wbar = wbar;
    MTR_DISPATCH(prec, k_mean, dim3(ew_grid(N)), dim3(256), 0, st, ma);
    return hipGetLastError();
}

hipError_t launch_residual(int prec, int64_t P, const void *p, const float *s, const float *c, void *r,
                           hipStream_t st)
{
    if (prec == MTR_PREC_BF16)
        hipLaunchKernelGGL(k_residual<__hip_bfloat16>, dim3(ew_grid(P)), dim3(256), 0, st, P,
                           (const __hip_bfloat16 *)p, s, c, (__hip_bfloat16 *)r);
    else if (prec == MTR_PREC_FP16)
        hipLaunchKernelGGL(k_residual<__half>, dim3(ew_grid(P)), dim3(256), 0, st, P, (const __half *)p,
                           s, c, (__half *)r);
    else
        hipLaunchKernelGGL(k_residual<float>, dim3(ew_grid(P)), dim3(256), 0, st, P, (const float *)p, s,
                           c, (float *)r);
    return hipGetLastError();
}

hipError_t launch_to_storage(int prec, int64_t n, const float *src, void *dst, hipStream_t st)
{
    if (prec == MTR_PREC_BF16)
        hipLaunchKernelGGL(k_to_storage<__hip_bfloat16>, dim3(ew_grid(n)), dim3(256), 0, st, n, src,
                           (__hip_bfloat16 *)dst);
    else if (prec == MTR_PREC_FP16)
        hipLaunchKernelGGL(k_to_storage<__half>, dim3(ew_grid(n)), dim3(256), 0, st, n, src, (__half *)dst);
    else
        hipLaunchKernelGGL(k_to_storage<float>, dim3(ew_grid(n)), dim3(256), 0, st, n, src, (float *)dst);
    return hipGetLastError();
}

hipError_t launch_from_storage(int prec, int64_t n, const void *src, float *dst, hipStream_t st)
{
    if (prec == MTR_PREC_BF16)
        hipLaunchKernelGGL(k_from_storage<__hip_bfloat16>, dim3(ew_grid(n)), dim3(256), 0, st, n,
                           (const __hip_bfloat16 *)src, dst);
    else if (prec == MTR_PREC_FP16)
        hipLaunchKernelGGL(k_from_storage<__half>, dim3(ew_grid(n)), dim3(256), 0, st, n, (const __half *)src,
                           dst);
    else
        hipLaunchKernelGGL(k_from_storage<float>, dim3(ew_grid(n)), dim3(256), 0, st, n, (const float *)src,
                           dst);
    return hipGetLastError();
}

hipError_t launch_fill(int prec, int64_t n, void *dst, float v, hipStream_t st)
{
    if (prec == MTR_PREC_BF16)
        hipLaunchKernelGGL(k_fill<__hip_bfloat16>, dim3(ew_grid(n)), dim3(256), 0, st, n, (__hip_bfloat16 *)dst, v);
    else if (prec == MTR_PREC_FP16)
        hipLaunchKernelGGL(k_fill<__half>, dim3(ew_grid(n)), dim3(256), 0, st, n, (__half *)dst, v);
    else
        hipLaunchKernelGGL(k_fill<float>, dim3(ew_grid(n)), dim3(256), 0, st, n, (float *)dst, v);
    return hipGetLastError();
}

hipError_t launch_permute(int prec, const void *src, void *dst, int nX, int nY, int nZ, const int64_t *ss,
                          const int64_t *ds, hipStream_t st)
{
    PermArgs pa;
    pa.n[0] = nX;
    pa.n[1] = nY;
    pa.n[2] = nZ;
    for (int q = 0; q < 3; q++) {
        pa.ss[q] = ss[q];
        pa.ds[q] = ds[q];
    }
    // Axes of extent 1 carry no layout: compare the others.
    bool same = true;
    for (int q = 0; q < 3; q++)
        if (pa.n[q] > 1 && ss[q] != ds[q]) same = false;
    if (same)
        return hipMemcpyAsync(dst, src, (size_t)nX * nY * nZ * (prec == MTR_PREC_FP32 ? 4 : 2),
                              hipMemcpyDeviceToDevice, st);
    pa.A = pa.B = -1;
    for (int q = 0; q < 3; q++) {
        if (pa.n[q] > 1 && ss[q] == 1 && pa.A < 0) pa.A = q;
        if (pa.n[q] > 1 && ds[q] == 1 && pa.B < 0) pa.B = q;
    }
    if (pa.A < 0 || pa.B < 0 || pa.A == pa.B) return hipErrorInvalidValue;
    pa.C = 3 - pa.A - pa.B;
    const int64_t nb = (int64_t)((pa.n[pa.A] + 31) / 32) * ((pa.n[pa.B] + 31) / 32) * pa.n[pa.C];
    if (nb > 0x7fffffff) return hipErrorInvalidValue;
    if (prec == MTR_PREC_BF16)
        hipLaunchKernelGGL(k_permute<__hip_bfloat16>, dim3((unsigned)nb), dim3(256), 0, st,
                           (const __hip_bfloat16 *)src, (__hip_bfloat16 *)dst, pa);
    else if (prec == MTR_PREC_FP16)
        hipLaunchKernelGGL(k_permute<__half>, dim3((unsigned)nb), dim3(256), 0, st, (const __half *)src, (__half *)dst,
                           pa);
    else
        hipLaunchKernelGGL(k_permute<float>, dim3((unsigned)nb), dim3(256), 0, st, (const float *)src, (float *)dst,
                           pa);
    return hipGetLastError();
}

int fp_col_blocks(const AgentDev &ag)
{
    return 8 * ((ag.nstrip + 7) / 8) * ((ag.A + kColWaves - 1) / kColWaves);
}

int fp_col_max_nu() { return (int)(65536 / (kColWaves * 2 * sizeof(float))) - 2; }

template <class T>
static hipError_t launch_fp_col_t(const AgentDev *d_ags, int nag, int max_blocks, int slots, bool unit,
                                  hipStream_t st)
{
    const size_t shm = (size_t)kColWaves * slots * sizeof(float2);
    dim3 grid(max_blocks, nag), block(64 * kColWaves);
    if (unit) hipLaunchKernelGGL((k_fp_col<T, true>), grid, block, shm, st, d_ags);
    else hipLaunchKernelGGL((k_fp_col<T, false>), grid, block, shm, st, d_ags);
    return hipGetLastError();
}

hipError_t launch_fp_col(int prec, const AgentDev *d_ags, int nag, int max_blocks, int slots, bool unit,
                         hipStream_t st)
{
    if (prec == MTR_PREC_BF16) return launch_fp_col_t<__hip_bfloat16>(d_ags, nag, max_blocks, slots, unit, st);
    if (prec == MTR_PREC_FP16) return launch_fp_col_t<__half>(d_ags, nag, max_blocks, slots, unit, st);
    return launch_fp_col_t<float>(d_ags, nag, max_blocks, slots, unit, st);
}

int fp_pfx_blocks(const AgentDev &ag)
{
    return 8 * ((ag.nstrip + 7) / 8) * ((ag.A + kPfxWaves - 1) / kPfxWaves);
}

int fp_pfx_lds(int col_win, int L, int nz, bool unit)
{
    const int slots = pfx_slots(col_win, L);
    return pfx_cols(slots, nz) > 0 || unit ? pfx_lds_bytes(slots, nz, unit) : 0;
}

template <class T>
static hipError_t launch_fp_pfx_t(const AgentDev *d_ags, int nag, int max_blocks, int lds, bool unit, bool win,
                                  hipStream_t st)
{
    dim3 grid(max_blocks, nag), block(64 * kPfxWaves);
    if (win) {
        if (unit) hipLaunchKernelGGL((k_fp_pfx<T, true, true>), grid, block, lds, st, d_ags);
        else hipLaunchKernelGGL((k_fp_pfx<T, false, true>), grid, block, lds, st, d_ags);
    } else {
        if (unit) hipLaunchKernelGGL((k_fp_pfx<T, true, false>), grid, block, lds, st, d_ags);
        else hipLaunchKernelGGL((k_fp_pfx<T, false, false>), grid, block, lds, st, d_ags);
    }
    return hipGetLastError();
}

hipError_t launch_fp_pfx(int prec, const AgentDev *d_ags, int nag, int max_blocks, int lds, bool unit, bool win,
                         hipStream_t st)
{
    if (lds <= 0 || lds > kPfxLds) return hipErrorInvalidValue;
    if (prec == MTR_PREC_BF16) return launch_fp_pfx_t<__hip_bfloat16>(d_ags, nag, max_blocks, lds, unit, win, st);
    if (prec == MTR_PREC_FP16) return launch_fp_pfx_t<__half>(d_ags, nag, max_blocks, lds, unit, win, st);
    return launch_fp_pfx_t<float>(d_ags, nag, max_blocks, lds, unit, win, st);
}

int fp_col_res_blocks(const AgentDev &ag)
{
    const int64_t P = (int64_t)ag.A * ag.nu * ag.nv;
    return (int)std::min<int64_t>((P + 255) / 256, 2048);
}

template <class T>
static hipError_t launch_fp_col_res_t(const AgentDev *d_ags, int nag, int blocks, int mode, float *c_out,
                                      hipStream_t st)
{
    dim3 grid(blocks, nag), block(256);
    if (mode == 2) hipLaunchKernelGGL((k_fp_col_res<T, 2>), grid, block, 0, st, d_ags, c_out);
    else hipLaunchKernelGGL((k_fp_col_res<T, 0>), grid, block, 0, st, d_ags, c_out);
    return hipGetLastError();
}

hipError_t launch_fp_col_res(int prec, const AgentDev *d_ags, int nag, int blocks, int mode, float *c_out,
                             hipStream_t st)
{
    if (prec == MTR_PREC_BF16) return launch_fp_col_res_t<__hip_bfloat16>(d_ags, nag, blocks, mode, c_out, st);
    if (prec == MTR_PREC_FP16) return launch_fp_col_res_t<__half>(d_ags, nag, blocks, mode, c_out, st);
    return launch_fp_col_res_t<float>(d_ags, nag, blocks, mode, c_out, st);
}

int bp_col_blocks(const AgentDev &ag, int rw)
{
    const int n1 = ag.col_p1y ? ag.Y : ag.X, n2 = ag.col_p1y ? ag.X : ag.Y;
    const int nb = ((n1 + 63) / 64) * ((n2 + kColBpWaves - 1) / kColBpWaves) * col_runs(ag.Zs, ag.Ze, rw);
    return ((nb + 7) / 8) * 8;
}

int bp_cs_blocks(const AgentDev &ag, int rw)
{
    const int ntz = col_runs(ag.Zs, ag.Ze, rw), nzg = (ntz + kCsWaves - 1) / kCsWaves;
    const int nb = ag.ngrp * nzg;
    return ((nb + 7) / 8) * 8;
}

int bp_cs_planes() { return kCsWaves * 16; }

template <class T>
static hipError_t launch_bp_cs_t(const AgentDev *d_ags, int nag, int max_blocks, int mode, float lambda,
                                 hipStream_t st)
{
    dim3 grid(max_blocks, nag), block(64 * kCsWaves);
    if (mode == 2) hipLaunchKernelGGL((k_bp_cs<T, 16, 2>), grid, block, 0, st, d_ags, lambda);
    else if (mode == 1) hipLaunchKernelGGL((k_bp_cs<T, 16, 1>), grid, block, 0, st, d_ags, lambda);
    else hipLaunchKernelGGL((k_bp_cs<T, 16, 0>), grid, block, 0, st, d_ags, lambda);
    return hipGetLastError();
}

hipError_t launch_bp_cs(int prec, const AgentDev *d_ags, int nag, int max_blocks, int mode, float lambda,
                        hipStream_t st)
{
    if (prec == MTR_PREC_BF16) return launch_bp_cs_t<__hip_bfloat16>(d_ags, nag, max_blocks, mode, lambda, st);
    if (prec == MTR_PREC_FP16) return launch_bp_cs_t<__half>(d_ags, nag, max_blocks, mode, lambda, st);
    return launch_bp_cs_t<float>(d_ags, nag, max_blocks, mode, lambda, st);
}

int bp_col_waves() { return kColBpWaves; }
int fp_col_waves() { return kColWaves; }

template <class T, int RW>
static void launch_bp_col_rw(dim3 grid, const AgentDev *d_ags, int mode, float lambda, hipStream_t st)
{
    const dim3 block(64 * kColBpWaves);
    if (mode == 2) hipLaunchKernelGGL((k_bp_col<T, RW, 2>), grid, block, 0, st, d_ags, lambda);
    else if (mode == 1) hipLaunchKernelGGL((k_bp_col<T, RW, 1>), grid, block, 0, st, d_ags, lambda);
    else hipLaunchKernelGGL((k_bp_col<T, RW, 0>), grid, block, 0, st, d_ags, lambda);
}

template <class T>
static hipError_t launch_bp_col_t(const AgentDev *d_ags, int nag, int max_blocks, int mode, float lambda, int rw,
                                  hipStream_t st)
{
    // runs of 16 planes (runs of 32 need more than 128 VGPRs per lane)
    if (rw != 16) return hipErrorInvalidValue;
    dim3 grid(max_blocks, nag);
    launch_bp_col_rw<T, 16>(grid, d_ags, mode, lambda, st);
    return hipGetLastError();
}

hipError_t launch_bp_col(int prec, const AgentDev *d_ags, int nag, int max_blocks, int mode, float lambda, int rw,
                         hipStream_t st)
{
    if (prec == MTR_PREC_BF16) return launch_bp_col_t<__hip_bfloat16>(d_ags, nag, max_blocks, mode, lambda, rw, st);
    if (prec == MTR_PREC_FP16) return launch_bp_col_t<__half>(d_ags, nag, max_blocks, mode, lambda, rw, st);
    return launch_bp_col_t<float>(d_ags, nag, max_blocks, mode, lambda, rw, st);
}

hipError_t launch_reduce(const double *part, int64_t n, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, st, part, n, out);
    return hipGetLastError();
}

hipError_t launch_plain_splat(const AgentDev &ag, float *s, float *c, hipStream_t st)
{
    dim3 block(8, 8, 8), grid((ag.X + 7) / 8, (ag.Y + 7) / 8, (ag.Z + 7) / 8);
    hipLaunchKernelGGL(k_plain_splat, grid, block, 0, st, ag, s, c);
    return hipGetLastError();
}

hipError_t launch_plain_bp(const AgentDev &ag, const float *r, float lambda, int bpt, hipStream_t st)
{
    dim3 block(8, 8, 8), grid((ag.X + 7) / 8, (ag.Y + 7) / 8, (ag.Z + 7) / 8);
    hipLaunchKernelGGL(k_plain_bp, grid, block, 0, st, ag, r, lambda, bpt);
    return hipGetLastError();
}

}  // namespace mtr
