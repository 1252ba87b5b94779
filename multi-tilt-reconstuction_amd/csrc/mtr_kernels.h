// Launchers of the HIP kernels (host-callable, defined in mtr_kernels.hip).
#pragma once

#include <vector>

#include "mtr_internal.h"

namespace mtr {

int fp_blocks(const AgentDev &ag, bool lanes_k);
int bp_blocks(const AgentDev &ag, int tz);
int consensus_blocks(int64_t N);
// Row-separable forward projection streaming one ray per wave over 64
// planes (plane axis contiguous, u independent of it), fused residual.
int fp_stream_blocks(const AgentDev &ag, int cols_per_wave);
hipError_t launch_fp_stream(int prec, const AgentDev *d_ags, int nag, int max_blocks, int nc, int cols_per_wave,
                            hipStream_t st);
// Agents streaming in Z segments (ag.nseg > 1): the segments' sums and the
// residual, grid-stride over the pixels (blocks as fp_col_res_blocks).
hipError_t launch_fp_seg_res(int prec, const AgentDev *d_ags, int nag, int blocks, hipStream_t st);
// Row-separable back-projection (runs along the plane axis W).
// rw: voxels per thread along W (8, 16 or 32).
int bp_sep_blocks(const AgentDev &ag, int W, int P1, int P2, int rw);
int bp_sep_pitch();
int bp_sep_waves();
// The window table of ag.sep_win for the views bsv[A] of an agent whose box
// and detector width ag holds, tiles along P1 and P2.
std::vector<int2> bp_sep_windows(const BpSepView *bsv, int A, const AgentDev &ag, int P1, int P2);
hipError_t launch_bp_sep(int prec, const AgentDev *d_ags, int nag, int max_blocks, bool bpt, float lambda, int W,
                         int P1, int P2, int rw, hipStream_t st);

// Column plan (v independent of Z and of the view; Z-contiguous volume).
// Forward projection into partial rows pm/pn (unit: x = 1), then the sum
// and residual (mode 0) or the normaliser into c_out (mode 2).
int fp_col_blocks(const AgentDev &ag);
int fp_col_max_nu();
hipError_t launch_fp_col(int prec, const AgentDev *d_ags, int nag, int max_blocks, int slots, bool unit,
                         hipStream_t st);
// The same forward projection from per-column prefix moments (8 views per
// block sharing each column's moments): lds = fp_pfx_lds(col_win, nu + 2, Z, unit)
// bytes of the batch's largest agent; 0 when a column's moments and the
// rows do not fit (use launch_fp_col then).
int fp_pfx_blocks(const AgentDev &ag);
int fp_pfx_lds(int col_win, int L, int nz, bool unit);
// win: every agent of the batch keeps a row window (col_win > 0), or none does.
hipError_t launch_fp_pfx(int prec, const AgentDev *d_ags, int nag, int max_blocks, int lds, bool unit, bool win,
                         hipStream_t st);
int fp_col_res_blocks(const AgentDev &ag);
hipError_t launch_fp_col_res(int prec, const AgentDev *d_ags, int nag, int blocks, int mode, float *c_out,
                             hipStream_t st);
// Back-projection with runs of rw = 16 voxels along Z; the staged
// footprint (rows x col_pitch pairs) must fit col_elems().  mode 0: SIRT
// update, 1: BPT, 2: the geometry-only normaliser C into ag.bpc (needed by
// modes 0 and 1).
int bp_col_blocks(const AgentDev &ag, int rw);
// The same back-projection on strip tiles (ag.col_grp, runs of 16 along
// Z, bp_cs_planes() planes per block); every group's footprint must span at
// most cs_max_width() detector columns in every view.
int bp_cs_blocks(const AgentDev &ag, int rw);
int bp_cs_planes();
int cs_max_width();
// The window table of ag.cs_win for groups grp[0 .. ngrp) of a box whose
// planes are [Zs, Ze), views xc[4 A].
std::vector<int2> bp_cs_windows(const ColGroup *grp, int ngrp, const double *xc, int A, int Zs, int Ze);
hipError_t launch_bp_cs(int prec, const AgentDev *d_ags, int nag, int max_blocks, int mode, float lambda,
                        hipStream_t st);
int bp_col_waves();
int fp_col_waves();
int col_elems();
hipError_t launch_bp_col(int prec, const AgentDev *d_ags, int nag, int max_blocks, int mode, float lambda, int rw,
                         hipStream_t st);

// Forward projection of nag agents (device table d_ags), mode 0 = fused
// residual, 1 = raw s/c into s_out/c_out, 2 = normaliser into c_out.
hipError_t launch_fp(int prec, const AgentDev *d_ags, int nag, int max_blocks, int mode, bool lanes_k,
                     float *s_out, float *c_out, hipStream_t st);
// Back-projection (bpt: initial model) of nag agents; src overrides the
// agents' residual buffers when non-null (single-agent use).
hipError_t launch_bp(int prec, const AgentDev *d_ags, int nag, int max_blocks, int tz, bool bpt,
                     float lambda, const void *src, int lds_cap, hipStream_t st);
hipError_t launch_consensus(int prec, int K, int64_t N, float rho, int write_next, void *const *w,
                            void *const *x, void *wbar, double *partial, int blocks, hipStream_t st);
hipError_t launch_init_sum(int prec, int K, int64_t N, void *const *w, void *const *x, void *wbar,
                           hipStream_t st);
hipError_t launch_form_z(int prec, int64_t N, const void *w, const void *wbar, void *z, hipStream_t st);
hipError_t launch_mann(int prec, int64_t N, const void *v, void *w, const void *z, float rho,
                       hipStream_t st);
hipError_t launch_mann_z(int prec, int64_t N, const void *v, void *w, const void *wbar, float rho,
                         hipStream_t st);
hipError_t launch_mean(int prec, int K, int64_t N, const void *const *w, void *wbar, hipStream_t st);
hipError_t launch_residual(int prec, int64_t P, const void *p, const float *s, const float *c, void *r,
                           hipStream_t st);
hipError_t launch_to_storage(int prec, int64_t n, const float *src, void *dst, hipStream_t st);
hipError_t launch_from_storage(int prec, int64_t n, const void *src, float *dst, hipStream_t st);
hipError_t launch_fill(int prec, int64_t n, void *dst, float v, hipStream_t st);
// dst (strides ds) = src (strides ss) for an nX x nY x nZ volume whose
// contiguous axes differ.
hipError_t launch_permute(int prec, const void *src, void *dst, int nX, int nY, int nZ, const int64_t *ss,
                          const int64_t *ds, hipStream_t st);
hipError_t launch_reduce(const double *part, int64_t n, double *out, hipStream_t st);
hipError_t launch_plain_splat(const AgentDev &ag, float *s, float *c, hipStream_t st);
hipError_t launch_plain_bp(const AgentDev &ag, const float *r, float lambda, int bpt, hipStream_t st);

}  // namespace mtr
