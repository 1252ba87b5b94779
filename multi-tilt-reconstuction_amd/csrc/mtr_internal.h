// Internal types shared by the HIP kernels and the host-side context.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#include "mtr.h"

namespace mtr {

constexpr int kMaxAgents = 16;  // MAX_TILT, ReadAndInitialOptions.h:17

// Per-view geometry record in HBM, read through scalar loads.
// c[0..3] = xc[4a..4a+3], c[4..7] = yc[4a..4a+3] (the reference's
// coefficient pairs, FunctionOnGPU.cu:17-18).
struct ViewCoef {
    double c[8];
};

// Ray-driven forward-projection plan of one view.  A view is
// "row-separable" when v depends on exactly one voxel axis W with
// coefficient +-1 and an integer offset: detector row k then collects
// exactly the voxels of plane W(k) (weight 1) and the forward projection of
// the row is a 2-D gather inside that plane.  Other views use the general
// 3-D plan: planes perpendicular to the driving axis D (largest component
// of the ray direction), and inside each plane the bounding box of the
// support parallelogram |u - i| < 1, |v - k| < 1.
struct FpPlan {
    int kind;      // 0 = general, 1 = row-separable
    int D, P, Q;   // general: driving axis and in-plane axes (0=X,1=Y,2=Z)
    int W, L, M;   // separable: plane axis, line (drive) axis, cross axis
    int wsign;     // separable: bW (+1/-1)
    int ncand;     // separable: lattice points per line, floor(2 Em) + 1
    int pfx_view;  // column-plan FP: the view that wave slot v of a block group takes
    double woff;   // separable: W = (k + 1 - yc0) * wsign
    // general: crossing point (p*, q*) = p0 + i*pi + k*pk + d*pd
    double p0, pi, pk, pd;
    double q0, qi, qk, qd;
    double Ep, Eq;  // half extents of the support bounding box
    // separable: m*(l) = m0 + i*mi + w*mw + l*ml ; half extent Em
    double m0, mi, mw, ml, Em;
};

// Per-view constants of the run-along-W back-projection (row-separable
// views with u independent of W): u = c0 + a1*p1 + a2*p2 over the in-plane
// axes (P1, P2); detector row of plane w = k0 + kstep*w; over a full tile
// (64 x waves voxels) u spans [u(origin) + umin, u(origin) + umax].
struct BpSepView {
    double c0, a1, a2, umin, umax;
    int k0, kstep;
};

// Back-projection tile of a column-plan agent: up to 64 consecutive
// columns of one strip (col_xy[j0 .. j0 + n)), all on detector rows k0 and
// k0 + 1, with the bounding box of their (X, Y).
struct ColGroup {
    int j0, n;
    int xmin, xmax, ymin, ymax;
    int k0, pad;
};

// Device-side description of one agent, passed by value to kernels.
struct AgentDev {
    int nu, nv, A;
    int Xs, Xe, Ys, Ye, Zs, Ze;  // volume box (absolute voxel coordinates)
    int X, Y, Z;
    // Element strides of detector row k and column i within one view of
    // p, c and r: (nu, 1) for the [A][nv][nu] layout, (1, nv) for the
    // row-fastest [A][nu][nv] layout of row-separable agents.
    int pk, pi;
    int64_t sx, sy, sz;    // element strides of the working volume x
    const ViewCoef *coef;  // [A]
    const FpPlan *plan;    // [A]
    const BpSepView *bsv;  // [A] (row-separable agents only)
    const void *p;         // projections (storage precision), strides pk/pi
    const float *c;        // FP normaliser (FP32 always), strides pk/pi
    void *r;               // residual, strides pk/pi
    void *x;               // working volume, strides sx/sy/sz
    double *partial;       // per-block FP64 partial sums
    // Column plan (v independent of Z and identical in every view; the
    // working volume is Z-contiguous): the (X, Y) columns grouped into
    // strips by k0 = floor(v), strip s holding k0 = s - 1.
    const int2 *col_xy;    // absolute (X, Y) of each column, strip by strip
    const float2 *col_w;   // row weights (1 - dv, dv) of each column, formed in FP64
    const int *strip;      // [nstrip + 1] first column of each strip
    int nstrip;            // nv + 1 (k0 = -1 .. nv - 1)
    int col_p1y;           // back-projection tiles: lanes along Y (1) or X (0)
    int col_pitch;         // back-projection LDS pitch (pairs per staged row)
    int col_win;           // FP row window in slots (power of two), 0: the whole width
    int col_jit;           // FP row window: drop of a column's first slot below an earlier column's
    int col_span;          // FP row window: a column's last slot past its first
    float *pm, *pn;        // FP partial rows [A][nv][nu]: from strip k and strip k - 1
    float *bpc;            // back-projection normaliser C of every voxel (working layout)
    const ColGroup *col_grp;  // strip-tiled back-projection: column groups
    int ngrp;
    // strip-tiled back-projection: detector-column window (first, last) of
    // each block's footprint in each view, [(group * nzg + zgroup) * A + view]
    const int2 *cs_win;
    // separable back-projection: the same per (P1, P2) tile and view,
    // [(tile1 + nt1 * tile2) * A + view]
    const int2 *sep_win;
    // Streaming forward projection in Z segments of seg_len planes (nseg >
    // 1): per-segment partial ray sums [nseg][A][...] (strides pk/pi),
    // added and turned into the residual by a second kernel.
    float *seg_part;
    int nseg, seg_len;
    // Streaming forward projection: per segment, the (view, column group)
    // pairs ordered by where their rays cross the segment's middle plane,
    // packed as view + A * group (null: view fastest)
    const int *seg_order;
};

}  // namespace mtr
