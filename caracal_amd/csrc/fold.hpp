// Folded step kernel for the reference normal-mode transform.
//
// The reference transforms with rfft/irfft, both the real part of the DFT
// over sqrt(N): y_k = N^-1/2 sum_j cos(2 pi k j / N) x_j (src/rfft.f90,
// src/irfft.f90, used both ways in src/verlet.f90:377-463).  Since
// cos(2 pi k j / N) = cos(2 pi (N-k) j / N) = cos(2 pi k (N-j) / N), the
// modes k and N-k coincide, they are propagated with the same frequency, and
// the beads j and N-j after the back transform coincide as well: every
// normal-mode step leaves a ring polymer that is symmetric under the bead
// reflection j -> N-j, and the forces of a step are always evaluated on such
// a configuration.  This kernel therefore carries the H+1 distinct beads
// j = 0..H (H = N/2) of each ring polymer, and evaluates H+1 forces per step
// instead of N:
//   load    x_j <- (x_j + x_{N-j}) / 2 (exact for a symmetric state; for the
//           first step of a run only the sum x_j + x_{N-j} enters the
//           transform, so nothing is lost),
//   forward Y_k = sum_{c<H} M[k][c] F_c + s (-1)^k F_H             (k < H)
//           Y_H = s (sum_{c<H} w_c (-1)^c F_c + (-1)^H F_H)
//   back    the same with M again (the matrix is its own fold partner),
//   store   bead j and its mirror N-j,
// with s = N^-1/2, w_0 = 1, w_c = 2 otherwise and M[r][c] = s w_c cos(2 pi
// r c / N).  Per step and column this is 2 H^2 multiply-adds instead of the
// 2 N^2 of the unfolded transform.
//
// One block holds TPB = (256/L) / (H+1) trajectories.  The H x H block of M
// runs on the matrix cores (FP64 MFMA 16x16x4) with one wave per column tile
// and every row band of that tile, so that the modes after the forward pass,
// propagated in the accumulators, are directly the B operand of the back
// transform: the modes never go through LDS.  Row and column H enter as a
// rank-1 start of the accumulators and as an alternating sum over the four
// lane groups of a column.
#pragma once
#include <cmath>
#include <type_traits>

#include "philox.hpp"
#include "step_common.hpp"

namespace cam {
namespace fold {

template <class S, int NB>
struct Fold {
    static constexpr int H = NB / 2;                // beads 0..H are distinct
    static constexpr int NH = H + 1;                // LDS stride of a column (odd)
    static constexpr int SPB = kBlock / S::L;       // bead slots per block
    static constexpr int TPB = SPB / NH;            // trajectories per block
    static constexpr int NCOL = TPB * 2 * S::NC;    // p and q columns of the block
    static constexpr int CT = (NCOL + 15) / 16;     // 16-column tiles
    static constexpr int RB = (H + 15) / 16;        // 16-row bands of M
    static constexpr int KK = (H + 3) / 4;          // 4-deep k steps of M
    static_assert(NB >= 4 && TPB >= 1, "a folded ring polymer must fit in one block");
};

// The value v, which the compiler must treat as unknown from here on: lane
// indices derived from it are recomputed where they are used instead of
// being computed once and held in registers across the whole step loop
// (through the surface evaluation, which needs those registers).  Emits no
// instruction.
template <typename V>
__device__ __forceinline__ V opaque(V v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Sum over the four lane groups lk = lane / 16 of one column.
__device__ __forceinline__ double sum_lane_groups(double v)
{
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

// Folded normal-mode step of the block's columns, in place in xs
// ([column][bead 0..H], column = 2*(tl*NC + coord) + (0: p, 1: q)); tf is
// the block's LDS copy of M with row stride H + 1 (the 16 rows a wave reads
// per k step fall in distinct banks).
// Operand layout of v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and receives D[l/16 + 4r][l%16] in accumulator element r.
// tid is threadIdx.x.
template <class S, int NB, bool CENT>
__device__ __forceinline__ void nm_fold(double *xs, const double *prop, const double *tf, double *cent, int tid)
{
    using F = Fold<S, NB>;
    constexpr int H = F::H, NH = F::NH, NCOL = F::NCOL, CT = F::CT, RB = F::RB, KK = F::KK, NC = S::NC;
    constexpr int PW = PropTab<S::NA>::W;
    constexpr int NWAVE = kBlock / 64;
    constexpr int G = CT >= 2 * NWAVE ? 2 : 1;                 // column tiles in flight per wave
    constexpr int NG = (CT + NWAVE * G - 1) / (NWAVE * G);
    constexpr double s = 1.0 / kSqrt<NB>();
    constexpr double sh = (H & 1) ? -s : s;                    // s (-1)^H
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int li = lane & 15, lk = lane >> 4;
    // Every row or column index this lane touches is lk + 4n, so (-1)^index = (-1)^lk.
    const double sg = (lk & 1) ? -1.0 : 1.0;
    double av[RB][KK];                                          // M[16 b + li][4 kk + lk]
#pragma unroll
    for (int b = 0; b < RB; ++b)
#pragma unroll
        for (int kk = 0; kk < KK; ++kk) {
            const int row = 16 * b + li, c = 4 * kk + lk;
            av[b][kk] = row < H && c < H ? tf[row * (H + 1) + c] : 0.0;
        }
#pragma unroll 1
    for (int gi = 0; gi < NG; ++gi) {
        // Every wave runs G tiles per round, so that no MFMA sits under a
        // branch; a tile past the last one reads zeros and stores nothing.
        int col[G];
        bool ok[G];
        double fh[G], alt[G];
        dbl4 acc[G][RB];
#pragma unroll
        for (int t = 0; t < G; ++t) {
            col[t] = (wave + NWAVE * (gi * G + t)) * 16 + li;
            ok[t] = col[t] < NCOL;
            fh[t] = ok[t] ? xs[col[t] * NH + H] : 0.0;
            alt[t] = 0.0;
#pragma unroll
            for (int b = 0; b < RB; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[t][b][r] = 16 * b + lk + 4 * r < H ? s * sg * fh[t] : 0.0;
        }
        // forward: Y = M F + s (-1)^k F_H; the alternating sum of Y_H on the side
#pragma unroll
        for (int kk = 0; kk < KK; ++kk)
#pragma unroll
            for (int t = 0; t < G; ++t) {
                const int c = 4 * kk + lk;
                const double bv = ok[t] && c < H ? xs[col[t] * NH + c] : 0.0;
                alt[t] = fma(c == 0 ? 1.0 : 2.0 * sg, bv, alt[t]);
#pragma unroll
                for (int b = 0; b < RB; ++b)
                    acc[t][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[b][kk], bv, acc[t][b], 0, 0, 0);
            }
        double yh[G];
#pragma unroll
        for (int t = 0; t < G; ++t) yh[t] = s * sum_lane_groups(alt[t]) + sh * fh[t];
        // free ring-polymer map of each mode on (p, q) (src/verlet.f90:434-447);
        // p and q of one coordinate sit in neighbouring lanes, swapped with DPP
#pragma unroll
        for (int t = 0; t < G; ++t) {
            const bool isq = (col[t] & 1) != 0;
            const int pair = (ok[t] ? col[t] : 0) >> 1;
            const int tl = pair / NC, at = (pair % NC) / 3;
            const double *pt = prop + tl * NH * PW;
#pragma unroll
            for (int b = 0; b < RB; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    // rows past H - 1 hold zeros, which any table row leaves zero
                    const int mode = min(16 * b + lk + 4 * r, H);
                    const double other = quad::dpp<quad::perm(1, 0, 3, 2)>(acc[t][b][r]);
                    const double *pr = pt + mode * PW;
                    const double cw = pr[0], cx = pr[1 + 2 * at + (isq ? 1 : 0)];
                    acc[t][b][r] = isq ? other * cx + acc[t][b][r] * cw : acc[t][b][r] * cw + other * cx;
                }
            {
                const double other = quad::dpp<quad::perm(1, 0, 3, 2)>(yh[t]);
                const double *pr = pt + H * PW;
                const double cw = pr[0], cx = pr[1 + 2 * at + (isq ? 1 : 0)];
                yh[t] = isq ? other * cx + yh[t] * cw : yh[t] * cw + other * cx;
            }
            // The centroid of the propagated beads is mode 0 of q over sqrt(N).
            if (CENT && isq && ok[t] && lk == 0) cent[pair] = acc[t][0][0] * s;
        }
        // back: x = M Z + s (-1)^j Z_H, with the modes in the accumulators as
        // the B operand (rows 4 kk .. 4 kk + 3 are element kk % 4 of band kk / 4)
        dbl4 out[G][RB];
        double alt2[G];
#pragma unroll
        for (int t = 0; t < G; ++t) {
            alt2[t] = 0.0;
#pragma unroll
            for (int b = 0; b < RB; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = 16 * b + lk + 4 * r;
                    out[t][b][r] = row < H ? s * sg * yh[t] : 0.0;
                    alt2[t] = fma(row == 0 ? 1.0 : row < H ? 2.0 * sg : 0.0, acc[t][b][r], alt2[t]);
                }
        }
#pragma unroll
        for (int kk = 0; kk < KK; ++kk)
#pragma unroll
            for (int t = 0; t < G; ++t) {
                const double bv = acc[t][kk / 4][kk % 4];
#pragma unroll
                for (int b = 0; b < RB; ++b)
                    out[t][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[b][kk], bv, out[t][b], 0, 0, 0);
            }
        // A wave reads and writes only its own columns, in program order.
#pragma unroll
        for (int t = 0; t < G; ++t) {
            const double xh = s * sum_lane_groups(alt2[t]) + sh * yh[t];
            if (ok[t]) {
#pragma unroll
                for (int b = 0; b < RB; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = 16 * b + lk + 4 * r;
                        if (row < H) xs[col[t] * NH + row] = out[t][b][r];
                    }
                if (lk == 0) xs[col[t] * NH + H] = xh;
            }
        }
    }
    __syncthreads();
}

// One launch of nsteps steps, ENS CAM_ENS_NVE or CAM_ENS_CHILD, with the same
// step as the unfolded step kernel: p -= dt/2 g; free ring-polymer
// propagation; [CHILD: centroid to cbuf]; forces; p -= dt/2 g.  WPE: waves
// per SIMD the compiler is asked to fit (its register budget, 512 / WPE).
// Between two steps a lane holds q, p and g of its slots; during the forces
// only the gradient is live, p and q are read back from the column buffer
// after them.  The launch covers trajectories a.tbegin .. a.tend - 1 (state
// arrays and the centroid buffer keep their whole-batch strides).  NBUF: LDS
// column buffers (1 or 2, below).
template <class S, int NB, typename T, int ENS, int WPE, int NBUF>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE))) void fold_step_kernel(StepArgs a)
{
    using F = Fold<S, NB>;
    constexpr int NC = S::NC, L = S::L, SL = S::SLOTS, H = F::H, NH = F::NH, TPB = F::TPB;
    constexpr int PW = PropTab<S::NA>::W;
    constexpr bool CHILD = ENS == CAM_ENS_CHILD;
    // Column buffers.  A lane stages into and reads back from only its own
    // positions (its slots' coordinates at its bead j), and the transform,
    // which reads and writes every column, ends with a barrier; so one
    // buffer needs no barrier before the next step's staging.  With one, the
    // one-lane-per-bead surfaces (15 ring polymers a block at 32 beads) take
    // 53-70 KB of LDS and two or three blocks share a CU; with two (90 KB),
    // one block per CU leaves registers for other kernels' blocks, which a
    // batch of several surfaces sharing the chip runs faster with.  The CH5
    // kernels always take two, the layout their registers were tuned with.
    static_assert(NBUF == 1 || NBUF == 2, "one or two column buffers");
    __shared__ __align__(16) double X[NBUF][F::NCOL * NH];
    __shared__ double tfs[H * (H + 1)];       // M, row stride H + 1
    __shared__ __align__(16) double cent[CHILD ? TPB * NC : 1];
    __shared__ double prop[TPB * NH * PW];

    // A lane's place: bead slot `slot` of the block, ring polymer tl of the
    // block, distinct bead j and its mirror jm, coordinate cc[s] of its slot s.
    struct Place {
        int r, tl, j, jm, traj;
        bool live, act;
        int cc[SL];
    };
    const auto place = [&](int tid) {
        Place w;
        const int slot = tid / L;
        w.r = tid % L;
        w.live = slot < TPB * NH;                      // slots past the last ring polymer idle along
        w.tl = w.live ? slot / NH : TPB - 1;
        w.j = w.live ? slot % NH : H;
        w.jm = w.j == 0 ? 0 : NB - w.j;                // mirror bead (H is its own)
        const int traj_raw = a.tbegin + blockIdx.x * TPB + w.tl;
        w.act = w.live && traj_raw < a.tend;
        w.traj = traj_raw < a.tend ? traj_raw : a.tend - 1;
#pragma unroll
        for (int s = 0; s < SL; ++s) w.cc[s] = S::coord(s, w.r);
        return w;
    };
    const int tid = threadIdx.x;
    double q[SL], p[SL], g[SL];
    {
        const Place w = place(tid);
        const size_t base = (size_t)w.traj * NC * NB;
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            const bool ok = S::valid(s, w.r);
            const size_t o1 = base + (size_t)w.cc[s] * NB + w.j, o2 = base + (size_t)w.cc[s] * NB + w.jm;
            q[s] = ok ? 0.5 * (a.q[o1] + a.q[o2]) : 0.0;
            p[s] = ok ? 0.5 * (a.p[o1] + a.p[o2]) : 0.0;
            g[s] = ok ? 0.5 * (a.g[o1] + a.g[o2]) : 0.0;
        }
    }
    // per-trajectory mode table of modes 0..H (src/verlet.f90:413-432), as PropTab lays it out
    for (int i = tid; i < TPB * NH; i += kBlock) {
        const int t = i / NH, mode = i % NH;
        const int tr = min(a.tbegin + (int)blockIdx.x * TPB + t, a.tend - 1);
        const double bt = a.beta[tr];
        const double wk = (2.0 / (bt / NB)) * sin(mode * (3.141592653589793 / NB));
        const double wt = wk * a.dt;
        const double cw = cos(wt), sw = sin(wt);
        double *pr = prop + i * PW;
        pr[0] = mode == 0 ? 1.0 : cw;
#pragma unroll
        for (int at = 0; at < S::NA; ++at) {
            const double m = a.mass[at];
            if (mode == 0) {
                pr[1 + 2 * at] = 0.0;
                pr[2 + 2 * at] = a.dt / m;
            } else {
                const double wm = wk * m;
                pr[1 + 2 * at] = -wm * sw;
                pr[2 + 2 * at] = sw / wm;
            }
        }
    }
    for (int i = tid; i < H * H; i += kBlock) tfs[(i / H) * (H + 1) + i % H] = a.tfold[i];
    const double hdt = 0.5 * a.dt;
    double vb = 0.0;
    for (int it = 0; it < a.nsteps; ++it) {
        const int ti = opaque(tid);
        const Place w = place(ti);
        double *xb = X[NBUF == 1 ? 0 : it & 1];
        double *xs = xb + w.tl * 2 * NC * NH;      // this ring polymer's columns
        // half kick (src/verlet.f90:216-218)
#pragma unroll
        for (int s = 0; s < SL; ++s) p[s] = p[s] - hdt * g[s];
        // bead -> modes, propagate, back (src/verlet.f90:377-463)
        if (w.live) {
#pragma unroll
            for (int s = 0; s < SL; ++s)
                if (S::valid(s, w.r)) {
                    xs[(2 * w.cc[s]) * NH + w.j] = p[s];
                    xs[(2 * w.cc[s] + 1) * NH + w.j] = q[s];
                }
        }
        __syncthreads();
        nm_fold<S, NB, CHILD>(xb, prop, tfs, cent, ti);
        // The column buffer keeps this lane's p and q until the lane itself
        // stages the next step into it.
#pragma unroll
        for (int s = 0; s < SL; ++s) q[s] = xs[(2 * w.cc[s] + 1) * NH + w.j];
        if constexpr (CHILD) {
            // centroids of this child step (src/get_centroid.f90:66-85); their xi
            // is evaluated after the launch by xi_flag_kernel
            for (int i = ti; i < TPB * NC; i += kBlock) {
                const int tr = a.tbegin + blockIdx.x * TPB + i / NC;
                if (tr < a.tend) a.cbuf[((size_t)it * a.ntraj + tr) * NC + i % NC] = cent[i];
            }
        }
        // forces of the distinct beads (src/verlet.f90:772-777)
        {
            double x[SL];
#pragma unroll
            for (int s = 0; s < SL; ++s) x[s] = q[s];
            S::template egrad<T>(x, vb, g);
        }
        // second half kick (src/verlet.f90:1060-1062)
#pragma unroll
        for (int s = 0; s < SL; ++s) {
            const int ci = opaque(2 * w.cc[s] * NH + w.j);     // a fresh read, not the q held before
            p[s] = xs[ci] - hdt * g[s];
            q[s] = xs[ci + NH];
        }
    }
    const Place w = place(opaque(tid));
    if (w.act) {
        const size_t base = (size_t)w.traj * NC * NB;
#pragma unroll
        for (int s = 0; s < SL; ++s)
            if (S::valid(s, w.r)) {
                const size_t o1 = base + (size_t)w.cc[s] * NB + w.j, o2 = base + (size_t)w.cc[s] * NB + w.jm;
                a.q[o1] = q[s];
                a.p[o1] = p[s];
                a.g[o1] = g[s];
                if (w.jm != w.j) {
                    a.q[o2] = q[s];
                    a.p[o2] = p[s];
                    a.g[o2] = g[s];
                }
            }
        if (w.r == 0) {
            a.vbead[(size_t)w.traj * NB + w.j] = vb;
            if (w.jm != w.j) a.vbead[(size_t)w.traj * NB + w.jm] = vb;
        }
    }
}

template <class S, typename T, int WPE, int NBUF>
step_fn pick_fold_w(int nb, int ens, int *tpb)
{
    switch (nb) {
#define CAM_FOLD_CASE(N)                                                                      \
    case N:                                                                                   \
        *tpb = Fold<S, N>::TPB;                                                               \
        return ens == CAM_ENS_NVE ? fold_step_kernel<S, N, T, CAM_ENS_NVE, WPE, NBUF>         \
                                  : fold_step_kernel<S, N, T, CAM_ENS_CHILD, WPE, NBUF>;
        CAM_FOLD_CASE(4)
        CAM_FOLD_CASE(6)
        CAM_FOLD_CASE(8)
        CAM_FOLD_CASE(12)
        CAM_FOLD_CASE(16)
        CAM_FOLD_CASE(24)
        CAM_FOLD_CASE(32)
        CAM_FOLD_CASE(48)
        CAM_FOLD_CASE(64)
#undef CAM_FOLD_CASE
    default:
        return nullptr;
    }
}

}  // namespace fold
}  // namespace cam
