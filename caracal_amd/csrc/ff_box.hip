// QMDFF-form periodic box on the device: energy kernel and its hand-written
// adjoint (forces, parameter gradients, and the force-matching contraction
// d(v . dE/dx)/dtheta), plus the ring-polymer box engine built on them.
//
// Functional form (the reference's ff_eg/ff_nonb terms, src/ff_eg.f90:87-299,
// src/ff_nonb.f90:198-394):
//   bonds    De (1 - exp(-a (r - r0)))^2
//   angles   k (theta - theta0)^2
//   Coulomb  q_i q_j (erfc(a r)/r - zp (r - rc))        r < rc   (Zahn)
//   disp.    -(C6/(r^6 + R0^6) + sr42 C6/(r^8 + R0^8))  r < rc   (BJ-damped D3 form)
//   rep.     zab exp(-alpha r)/r                        r < rc
// Nonbonded pairs are intermolecular only; distances use the minimum image.
// theta = (De, a, r0, k, theta0, q_O, q_H, C6_OO, C6_OH, C6_HH).
//
// Layout: positions and gradients [bead][atom][3] (the reference's
// q(3,natoms,nbeads)).  The nonbonded kernel tiles j-atoms through LDS
// (64 per tile) and splits the j range over several blocks; every partial
// sum is reduced in a fixed order, so results are bitwise reproducible.
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "erfc_tables.hpp"
#include "common.hpp"
#include "nhc.hpp"
#include "philox.hpp"
#include "qmdff.hpp"

namespace cam {
namespace {

constexpr int kNP = 10;       // parameters
constexpr int kNE = 1 + 2 * kNP;   // energy + dE/dtheta + d(v.g)/dtheta
constexpr int kTile = 64;

struct FFDev {
    int natoms, nbonds, nangles;
    double box, rc, za, zp;
    double exmax;                  // a rc, bounded to [1e-3, 6]: erfc's argument range
    double sr42[2][2], r0bj[2][2], zrep[2][2], arep[2][2];
    const int *type, *mol, *bonds, *angles, *csr_ptr, *csr_term;
};

struct Theta {
    double p[kNP];
};

// Kahan-compensated accumulator for FP32 terms; plain FP64 otherwise.
template <typename T>
struct Acc {
    double s = 0.0;
    __device__ void add(T v) { s += (double)v; }
    __device__ double get() const { return s; }
};
template <>
struct Acc<float> {
    float s = 0.f, c = 0.f;
    __device__ void add(float v)
    {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ double get() const { return (double)s - (double)c; }
};

template <typename T>
__device__ __forceinline__ T min_image(T d, T box) { return d - box * rint(d / box); }

__device__ __forceinline__ float erfc_t(float x) { return erfcf(x); }
__device__ __forceinline__ double erfc_t(double x) { return erfc(x); }
__device__ __forceinline__ float exp_t(float x) { return expf(x); }
__device__ __forceinline__ double exp_t(double x) { return exp(x); }
__device__ __forceinline__ float sqrt_t(float x) { return sqrtf(x); }
__device__ __forceinline__ double sqrt_t(double x) { return sqrt(x); }

// ---------------------------------------------------------------------------
// Pair structure of the box.  The atoms are kept in a spatial order: cells of
// the bead centroids, the atoms of a cell in index order.  It is rebuilt only
// when some bead has moved more than half the skin since the last build, or
// the bead count changed.  A row per sorted atom lists, in ascending sorted
// index, every atom j of another molecule whose bead centroid lies within
// rc + skin + margin + dev_i + dev_j of atom i's (dev: the largest distance of
// a bead from its centroid) at the build, so it holds every pair closer than
// rc + skin + margin in some bead; while no bead has moved more than skin/2,
// no pair left out comes closer than rc + margin in any bead.  The pair
// kernel runs a brick of 256 sorted atoms per block against LDS chunks of up
// to 512 sorted atoms (only the 128-atom tiles that some row of the brick
// reaches); in all-pairs mode every atom of every tile.  Both meet the pairs
// inside the cutoff in the same order, so both give the same bits.
// ---------------------------------------------------------------------------
constexpr int kBrick = 256;
constexpr int kPTile = 128;
constexpr int kChunkTiles = 4;              // tiles staged together
constexpr int kMaxCellsDim = 20;            // cells per dimension (8000 cells: 32 KB of LDS in the sort)
constexpr int kMaxSplit = 32;
constexpr double kSkin = 1.2;               // bohr
constexpr double kMargin = 1e-3;            // bohr: binary32 rounding of the mixed kernel's distances
enum { kCtlGo = 0, kCtlBuiltNb = 1, kCtlNc = 2, kCtlWords = 4 };

struct PairDev {
    int n, nb, nsplit, ntiles, rowcap, maskw, nbricks, cg;
    double box, rlist;
    const int *type, *mol;                  // original order
    int *perm;                              // sorted -> original index
    int *scell, *smt, *tcell;               // per sorted atom: cell, mol << 1 | type; per original atom: cell
    double *scen;                           // per sorted atom: centroid (3) and spread
    int *cstart, *ccount;                   // per cell: first sorted atom and count
    int *row;                               // [sorted atom][rowcap]: sorted j, ascending
    int *sst;                               // [sorted atom][nsplit + 1]: row offsets of the splits; -1: overflow
    unsigned *mask;                         // [brick][maskw]: tiles its rows reach
    double *xref;                           // [bead][atom][3]: positions at the last build
    double *cen;                            // per atom: wrapped bead centroid (3) and spread
    double *bdev;                           // per check block: largest spread
    int *bover;                             // per check block: some bead moved more than skin/2
    int *ctl;                               // go, bead count of the last build, cells per dimension
};

__device__ __forceinline__ int cell_of(double c, double L, int ncd)
{
    return min(max((int)(c / L * ncd), 0), ncd - 1);
}

// Every call: per atom whether some bead moved more than skin/2 since the
// last build, the bead centroid (the mean of the minimum-image offsets from
// bead 0, wrapped into the box) and the spread (largest bead distance from
// it); per block the flag and the largest spread.  A group of G lanes per
// atom (G a power of two dividing 64), lane sub taking beads sub, sub + G, ...
constexpr int kCheckBeads = 4;                       // beads per lane kept in registers

__global__ __launch_bounds__(256) void pl_check_kernel(PairDev p, const double *__restrict__ x)
{
    __shared__ int s_ow[4];
    __shared__ double s_dw[4];
    const int n = p.n, nb = p.nb, tid = threadIdx.x;
    const int G = p.cg, sub = tid & (G - 1);
    const int i = blockIdx.x * (256 / G) + tid / G;
    const double L = p.box;
    const bool act = i < n;
    const bool held = nb <= kCheckBeads * G;         // every bead of the lane in registers
    double x0[3] = { 0.0, 0.0, 0.0 }, sm[3] = { 0.0, 0.0, 0.0 }, v[kCheckBeads][3] = {};
    int over = 0;
    auto first = [&](int b, double *xv) {
        double q2 = 0.0;
        for (int c = 0; c < 3; ++c) {
            const size_t k = ((size_t)b * n + i) * 3 + c;
            xv[c] = x[k];
            const double e = min_image(xv[c] - p.xref[k], L);
            q2 += e * e;
            sm[c] += min_image(xv[c] - x0[c], L);
        }
        over |= !(q2 <= 0.25 * kSkin * kSkin);       // also NaN
    };
    if (act) {
        for (int c = 0; c < 3; ++c) x0[c] = x[3 * i + c];
        if (held) {
#pragma unroll
            for (int u = 0; u < kCheckBeads; ++u)
                if (sub + u * G < nb) first(sub + u * G, v[u]);
        } else {
            double xv[3];
            for (int b = sub; b < nb; b += G) first(b, xv);
        }
    }
    for (int o = G >> 1; o > 0; o >>= 1)
        for (int c = 0; c < 3; ++c) sm[c] += __shfl_xor(sm[c], o);
    double cc[3];
    for (int c = 0; c < 3; ++c) cc[c] = x0[c] + sm[c] / nb;
    double m2 = 0.0;
    auto second = [&](const double *xv) {
        double r2 = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double d = min_image(xv[c] - cc[c], L);
            r2 += d * d;
        }
        m2 = fmax(m2, r2);
    };
    if (act) {
        if (held) {
#pragma unroll
            for (int u = 0; u < kCheckBeads; ++u)
                if (sub + u * G < nb) second(v[u]);
        } else {
            for (int b = sub; b < nb; b += G) {
                double xv[3];
                for (int c = 0; c < 3; ++c) xv[c] = x[((size_t)b * n + i) * 3 + c];
                second(xv);
            }
        }
    }
    for (int o = G >> 1; o > 0; o >>= 1) m2 = fmax(m2, __shfl_xor(m2, o));
    double devi = act ? sqrt(m2) : 0.0;
    if (act && sub == 0) {
        for (int c = 0; c < 3; ++c) p.cen[4 * i + c] = cc[c] - L * floor(cc[c] / L);
        p.cen[4 * i + 3] = devi;
    }
    for (int o = 32; o > 0; o >>= 1) {
        over |= __shfl_xor(over, o);
        devi = fmax(devi, __shfl_xor(devi, o));
    }
    if ((tid & 63) == 0) {
        s_ow[tid >> 6] = over;
        s_dw[tid >> 6] = devi;
    }
    __syncthreads();
    if (tid == 0) {
        p.bover[blockIdx.x] = s_ow[0] | s_ow[1] | s_ow[2] | s_ow[3];
        p.bdev[blockIdx.x] = fmax(fmax(s_dw[0], s_dw[1]), fmax(s_dw[2], s_dw[3]));
    }
}

// One block: decides whether to rebuild (a flag of the check, or a new bead
// count) and, if so, orders the atoms by cell of their centroids (the atoms
// of a cell in index order).
constexpr int kSortThreads = 1024;

__global__ __launch_bounds__(kSortThreads) void pl_sort_kernel(PairDev p, int nchk)
{
    __shared__ int s_cnt[kMaxCellsDim * kMaxCellsDim * kMaxCellsDim];
    __shared__ int s_c[kSortThreads];
    __shared__ int s_ow[kSortThreads / 64];
    __shared__ double s_dw[kSortThreads / 64];
    const int n = p.n, nb = p.nb, tid = threadIdx.x;
    const double L = p.box;
    int go = 0;
    double dmax = 0.0;
    for (int k = tid; k < nchk; k += kSortThreads) {
        go |= p.bover[k];
        dmax = fmax(dmax, p.bdev[k]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        go |= __shfl_xor(go, o);
        dmax = fmax(dmax, __shfl_xor(dmax, o));
    }
    if ((tid & 63) == 0) {
        s_ow[tid >> 6] = go;
        s_dw[tid >> 6] = dmax;
    }
    __syncthreads();
    go = p.ctl[kCtlBuiltNb] != nb;
    dmax = 0.0;
    for (int w = 0; w < kSortThreads / 64; ++w) {
        go |= s_ow[w];
        dmax = fmax(dmax, s_dw[w]);
    }
    __syncthreads();
    if (tid == 0) {
        p.ctl[kCtlGo] = go;
        if (go) p.ctl[kCtlBuiltNb] = nb;
    }
    if (!go) return;
    // cells at least rc + skin + margin + twice the largest spread wide: the
    // atoms a row can hold lie in the 27 cells around the atom's own
    int ncd = (int)floor(L / (p.rlist + 2.0 * dmax));
    ncd = ncd < 3 ? 1 : min(ncd, kMaxCellsDim);
    const int nc3 = ncd * ncd * ncd;
    if (tid == 0) p.ctl[kCtlNc] = ncd;
    for (int c = tid; c < nc3; c += kSortThreads) s_cnt[c] = 0;
    for (int k = tid; k < p.nbricks * p.maskw; k += kSortThreads) p.mask[k] = 0u;
    __syncthreads();
    for (int a = tid; a < n; a += kSortThreads) {
        const double *cc = p.cen + 4 * a;
        const int c = (cell_of(cc[0], L, ncd) * ncd + cell_of(cc[1], L, ncd)) * ncd + cell_of(cc[2], L, ncd);
        p.tcell[a] = c;
        atomicAdd(&s_cnt[c], 1);
    }
    __syncthreads();
    if (tid == 0) {                                   // exclusive scan: first sorted atom of each cell
        int acc = 0;
        for (int c = 0; c < nc3; ++c) {
            const int k = s_cnt[c];
            p.cstart[c] = acc;
            p.ccount[c] = k;
            s_cnt[c] = acc;
            acc += k;
        }
    }
    __syncthreads();
    for (int base = 0; base < n; base += kSortThreads) {   // atoms in index order, a block at a time
        const int a = base + tid;
        const int c = a < n ? p.tcell[a] : -1;
        s_c[tid] = c;
        __syncthreads();
        if (a < n) {
            int r = 0;
            for (int t = 0; t < tid; ++t) r += s_c[t] == c;
            p.perm[s_cnt[c] + r] = a;
        }
        __syncthreads();
        if (a < n) atomicAdd(&s_cnt[c], 1);
        __syncthreads();
    }
    for (int s = tid; s < n; s += kSortThreads) {
        const int a = p.perm[s];
        p.smt[s] = p.mol[a] << 1 | p.type[a];
        p.scell[s] = p.tcell[a];
        for (int c = 0; c < 4; ++c) p.scen[4 * s + c] = p.cen[4 * a + c];
    }
}

// On a rebuild, one wave per sorted atom s (pl_rows_kernel's waves take the
// atoms in turn).  Copies its bead positions as the
// new reference, walks the 27 cells around its own in ascending order (so the
// row comes out in ascending sorted index), keeps every atom of another
// molecule within the list radius, marks the tiles the row reaches for its
// brick and records where each split's part of the row starts.
__device__ void pl_row(const PairDev &p, const double *__restrict__ x, int s, int lane)
{
    const int n = p.n, nb = p.nb;
    const int i = p.perm[s], ms = p.smt[s] >> 1;
    for (int k = lane; k < 3 * nb; k += 64) {
        const size_t o = ((size_t)(k / 3) * n + i) * 3 + k % 3;
        p.xref[o] = x[o];
    }
    const double L = p.box, rl = p.rlist;
    const double ci0 = p.scen[4 * s], ci1 = p.scen[4 * s + 1], ci2 = p.scen[4 * s + 2], di = p.scen[4 * s + 3];
    const int ncd = p.ctl[kCtlNc];
    int cl[3][3], ncl[3];
    {
        const int cs = p.scell[s];
        const int own[3] = { cs / (ncd * ncd), (cs / ncd) % ncd, cs % ncd };
        for (int d = 0; d < 3; ++d) {
            if (ncd == 1) {
                ncl[d] = 1;
                cl[d][0] = 0;
            } else {                                  // ncd >= 3: three distinct cells, ascending
                int v0 = (own[d] + ncd - 1) % ncd, v1 = own[d], v2 = (own[d] + 1) % ncd, t;
                if (v0 > v1) { t = v0; v0 = v1; v1 = t; }
                if (v1 > v2) { t = v1; v1 = v2; v2 = t; }
                if (v0 > v1) { t = v0; v0 = v1; v1 = t; }
                ncl[d] = 3;
                cl[d][0] = v0;
                cl[d][1] = v1;
                cl[d][2] = v2;
            }
        }
    }
    const int S = p.nsplit, T = p.ntiles;
    int *row = p.row + (size_t)s * p.rowcap;
    int *sst = p.sst + (size_t)s * (S + 1);
    unsigned *mrow = p.mask + (size_t)(s / kBrick) * p.maskw;
    int cnt = 0, nsp = 0;
    int mw = -1;                                      // mask word being gathered (tiles ascend with j)
    unsigned mbits = 0u;
    auto mark = [&](int t) {
        if ((t >> 5) != mw) {
            if (mw >= 0 && lane == 0) atomicOr(&mrow[mw], mbits);
            mw = t >> 5;
            mbits = 0u;
        }
        mbits |= 1u << (t & 31);
    };
    auto split_begin = [&](int sp) { return (int)((long long)sp * T / S) * kPTile; };
    for (int ax = 0; ax < ncl[0]; ++ax)
        for (int ay = 0; ay < ncl[1]; ++ay)
            for (int az = 0; az < ncl[2]; ++az) {
                const int c = (cl[0][ax] * ncd + cl[1][ay]) * ncd + cl[2][az];
                const int a0 = p.cstart[c], a1 = a0 + p.ccount[c];
                // candidate data of the next batch loaded ahead of its use
                int jq = a0 + lane;
                double q0 = 0.0, q1 = 0.0, q2 = 0.0, qd = 0.0;
                int qm = ms;
                if (jq < a1) {
                    q0 = p.scen[4 * jq];
                    q1 = p.scen[4 * jq + 1];
                    q2 = p.scen[4 * jq + 2];
                    qd = p.scen[4 * jq + 3];
                    qm = p.smt[jq] >> 1;
                }
                for (int j0 = a0; j0 < a1; j0 += 64) {
                    const int j = jq;
                    const double c0 = q0, c1 = q1, c2 = q2, cd = qd;
                    const int mj = qm;
                    jq = j0 + 64 + lane;
                    if (jq < a1) {
                        q0 = p.scen[4 * jq];
                        q1 = p.scen[4 * jq + 1];
                        q2 = p.scen[4 * jq + 2];
                        qd = p.scen[4 * jq + 3];
                        qm = p.smt[jq] >> 1;
                    }
                    bool keep = false;
                    if (j < a1 && j != s && mj != ms) {
                        const double d0 = min_image(ci0 - c0, L), d1 = min_image(ci1 - c1, L),
                                     d2 = min_image(ci2 - c2, L);
                        const double lim = rl + di + cd;
                        keep = d0 * d0 + d1 * d1 + d2 * d2 < lim * lim;
                    }
                    const unsigned long long m = __ballot(keep);
                    const int len = min(64, a1 - j0);
                    // split starts at or before this batch, then inside it
                    while (nsp <= S && split_begin(nsp) <= j0) {
                        if (lane == 0) sst[nsp] = cnt;
                        ++nsp;
                    }
                    while (nsp <= S && split_begin(nsp) < j0 + len) {
                        const int below = split_begin(nsp) - j0;
                        if (lane == 0) sst[nsp] = cnt + __popcll(m & ((1ull << below) - 1ull));
                        ++nsp;
                    }
                    if (keep) {
                        const int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
                        if (pos < p.rowcap) row[pos] = j;
                    }
                    if (m) {
                        mark((j0 + __builtin_ctzll(m)) / kPTile);
                        mark((j0 + 63 - __builtin_clzll(m)) / kPTile);
                    }
                    cnt += __popcll(m);
                }
            }
    while (nsp <= S) {
        if (lane == 0) sst[nsp] = cnt;
        ++nsp;
    }
    if (mw >= 0 && lane == 0) atomicOr(&mrow[mw], mbits);
    if (cnt > p.rowcap) {                             // row too long: this atom walks whole tiles
        for (int k = lane; k < p.maskw; k += 64) atomicOr(&mrow[k], 0xFFFFFFFFu);
        if (lane == 0) sst[0] = -1;
    }
}

__global__ __launch_bounds__(256) void pl_rows_kernel(PairDev p, const double *__restrict__ x)
{
    if (p.ctl[kCtlGo] == 0) return;
    const int lane = threadIdx.x & 63;
    for (int s = blockIdx.x * 4 + (threadIdx.x >> 6); s < p.n; s += gridDim.x * 4)
        pl_row(p, x, s, lane);
}

// erfc(x) for x >= 0 as ex2 = exp(-x^2) times the polynomial of erfcx(x) =
// exp(x^2) erfc(x) on [0, X] (erfc_tables.hpp; Horner in t = 2x/X - 1), X =
// 2.5 where a rc <= 2.5 and 6 beyond (WIDE); past X the polynomial is held at
// X, where erfc is below 2.2e-17.  The coefficients are compile-time
// constants: loaded from the kernel arguments, they were held in SGPRs across
// the pair loop and, spilled to VGPR lanes, read back with a VALU
// instruction each.
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double rint_t(double a) { return __builtin_rint(a); }
__device__ __forceinline__ float rint_t(float a) { return __builtin_rintf(a); }

// The FP64 coefficients are read from LDS (sc: a block's copy of the table)
// through an address the compiler must treat as new at every use: one LDS
// read per coefficient next to the polynomial, into a register the FMA then
// accumulates in.  As constants they were held across the pair loop (in
// SGPRs spilled to VGPR lanes, or in VGPRs) or copied into a VGPR one 32-bit
// half at a time for every use (a 64-bit constant has no inline form in a
// VOP3 instruction).
typedef __attribute__((address_space(3))) const double lds_double;

__device__ __forceinline__ lds_double *opaque_lds(lds_double *p)
{
    asm volatile("" : "+v"(p));
    return p;
}

template <typename T, bool WIDE>
__device__ __forceinline__ T erfc_series(T x, T ex2, lds_double *sc)
{
#pragma clang fp contract(off)
    namespace E = erfc_tab;
    constexpr double X = WIDE ? E::kWideX : E::kNarrowX;
    const T t = fma_t(min(x, (T)X), (T)(2.0 / X), (T)-1);
    if constexpr (sizeof(T) == 8) {
        // four interleaved Horner chains in v = t^4 (a quarter of the
        // dependent FMAs of one chain): p = (a + t b) + t^2 (c + t d)
        constexpr int NT = WIDE ? 36 : 24;
        static_assert(NT % 4 == 0, "four chains");
        lds_double *c = opaque_lds(sc);
        const T u = t * t, v = u * u;
        T h0 = c[NT - 4], h1 = c[NT - 3], h2 = c[NT - 2], h3 = c[NT - 1];
#pragma unroll
        for (int k = NT - 8; k >= 0; k -= 4) {
            h0 = fma_t(h0, v, (T)c[k]);
            h1 = fma_t(h1, v, (T)c[k + 1]);
            h2 = fma_t(h2, v, (T)c[k + 2]);
            h3 = fma_t(h3, v, (T)c[k + 3]);
        }
        return ex2 * fma_t(fma_t(h3, t, h2), u, fma_t(h1, t, h0));
    } else {
        constexpr int NT = WIDE ? 24 : 16;
        const float *c = WIDE ? E::kWide32 : E::kNarrow32;
        T h = c[NT - 1];
#pragma unroll
        for (int k = NT - 2; k >= 0; --k) h = fma_t(h, t, (T)c[k]);
        return ex2 * h;
    }
}

// Pair-loop arithmetic.  FP64: reciprocal and reciprocal square root from
// the hardware estimates with two Newton steps each (positive normal
// arguments only); binary32: the hardware instructions as they are (1 ulp).
__device__ __forceinline__ double rcp_pair(double a)
{
#pragma clang fp contract(off)
    double y = __builtin_amdgcn_rcp(a);
    y = y * fma(-a, y, 2.0);
    return y * fma(-a, y, 2.0);
}
__device__ __forceinline__ float rcp_pair(float a) { return __builtin_amdgcn_rcpf(a); }
__device__ __forceinline__ double rsq_pair(double a)
{
#pragma clang fp contract(off)
    double y = __builtin_amdgcn_rsq(a);
    const double h = 0.5 * a;
    y = y * fma(-h * y, y, 1.5);
    return y * fma(-h * y, y, 1.5);
}
__device__ __forceinline__ float rsq_pair(float a) { return __builtin_amdgcn_rsqf(a); }
__device__ __forceinline__ double exp_pair(double a) { return exp(a); }
__device__ __forceinline__ float exp_pair(float a) { return __expf(a); }

// Nonbonded pairs of a brick of sorted atoms (one lane each) against the
// tiles of split blockIdx.y; GRAD: gradient of the lane's atom, PAR:
// parameter derivatives (and with v, the force-matching contraction).
// LIST: the pairs of the rows; otherwise every atom of every tile.
template <typename T, bool GRAD, bool PAR, bool LIST, bool WIDE>
__device__ __forceinline__ void nonbonded_body(const FFDev &f, const PairDev &pl, const Theta &th,
                                               const double *__restrict__ x, const double *__restrict__ v,
                                               double *__restrict__ gpart, double *__restrict__ epart)
{
    // The arithmetic is written out operation by operation (fused only
    // where fma_t says so): the row walk and the all-pairs walk, compiled
    // apart, round every pair the same way.
#pragma clang fp contract(off)
    constexpr int kC = kChunkTiles * kPTile;
    __shared__ T sx[kC][3];
    __shared__ T sv[PAR ? kC : 1][3];
    __shared__ int smt[kC];
    __shared__ double red[kBrick / 64][11];
    __shared__ double scoef[36];                  // FP64 erfcx coefficients (erfc_series)
    if constexpr (sizeof(T) == 8) {
        const double *tab = WIDE ? erfc_tab::kWide64 : erfc_tab::kNarrow64;
        if (threadIdx.x < (WIDE ? 36 : 24)) scoef[threadIdx.x] = tab[threadIdx.x];
    }
    const int n = f.natoms;
    const int tid = threadIdx.x;
    const int brick = blockIdx.x, split = blockIdx.y, b = blockIdx.z, S = gridDim.y;
    const int s = brick * kBrick + tid;
    const bool act = s < n;
    const int i = act ? pl.perm[s] : 0;
    const double *xb = x + (size_t)b * n * 3;
    const double *vb = v ? v + (size_t)b * n * 3 : nullptr;
    const T xi0 = (T)xb[3 * i], xi1 = (T)xb[3 * i + 1], xi2 = (T)xb[3 * i + 2];
    T vi0 = 0, vi1 = 0, vi2 = 0;
    if (PAR && vb) {
        vi0 = (T)vb[3 * i];
        vi1 = (T)vb[3 * i + 1];
        vi2 = (T)vb[3 * i + 2];
    }
    const int mti = act ? pl.smt[s] : 0;
    const int ti = mti & 1, mi = mti >> 1;
    const T box = (T)f.box, ibox = (T)(1.0 / f.box), rc2 = (T)(f.rc * f.rc), za = (T)f.za, zp = (T)f.zp,
            rc = (T)f.rc;
    const T two_over_spi = (T)1.1283791670955126;
    const T qi = (T)(ti == 0 ? th.p[5] : th.p[6]);
    // per-partner-type constants, selected by the partner's type below (no
    // run-time indexing of parameter arrays)
    // per-partner-type constants (partner type 0 / 1), selected per pair
    const T qt0 = (T)th.p[5], qt1 = (T)th.p[6];
    const T c6t0 = (T)(ti == 0 ? th.p[7] : th.p[8]), c6t1 = (T)(ti == 0 ? th.p[8] : th.p[9]);
    const T r0t0 = (T)(ti == 0 ? f.r0bj[0][0] : f.r0bj[1][0]), r0t1 = (T)(ti == 0 ? f.r0bj[0][1] : f.r0bj[1][1]);
    const T srt0 = (T)(ti == 0 ? f.sr42[0][0] : f.sr42[1][0]), srt1 = (T)(ti == 0 ? f.sr42[0][1] : f.sr42[1][1]);
    const T zrt0 = (T)(ti == 0 ? f.zrep[0][0] : f.zrep[1][0]), zrt1 = (T)(ti == 0 ? f.zrep[0][1] : f.zrep[1][1]);
    const T alt0 = (T)(ti == 0 ? f.arep[0][0] : f.arep[1][0]), alt1 = (T)(ti == 0 ? f.arep[0][1] : f.arep[1][1]);
    Acc<T> g0, g1, g2;
    double e = 0.0;
    // parameter derivatives: q_O, q_H, C6 of (ti, partner) pairs OO/OH/HH;
    // dv* the same for the force-matching contraction (named scalars, no
    // run-time indexed arrays)
    double dq0 = 0.0, dq1 = 0.0, dc0 = 0.0, dc1 = 0.0, dc2 = 0.0;
    double dvq0 = 0.0, dvq1 = 0.0, dvc0 = 0.0, dvc1 = 0.0, dvc2 = 0.0;

    // minimum-image separation of i from the atom at LDS slot jj
    auto sep = [&](int jj, T &d0, T &d1, T &d2) -> T {
        d0 = xi0 - sx[jj][0];
        d1 = xi1 - sx[jj][1];
        d2 = xi2 - sx[jj][2];
        d0 = fma_t(-box, rint_t(d0 * ibox), d0);
        d1 = fma_t(-box, rint_t(d1 * ibox), d1);
        d2 = fma_t(-box, rint_t(d2 * ibox), d2);
        return fma_t(d0, d0, fma_t(d1, d1, d2 * d2));
    };
    // one pair inside the cutoff: j at LDS slot jj, separation d, r2 = |d|^2
    auto pair = [&](int jj, T d0, T d1, T d2, T r2) {
            const int tj = smt[jj] & 1;
            const bool jo = tj == 0;
            const T rinv = rsq_pair(r2);
            const T r = r2 * rinv;
            const T qj = jo ? qt0 : qt1;
            const T c6 = jo ? c6t0 : c6t1;
            const T xa = za * r;
            const T ex2 = exp_pair(-(xa * xa));
            const T ev = erfc_series<T, WIDE>(xa, ex2, (lds_double *)scoef);
            const T R0 = jo ? r0t0 : r0t1, sr = jo ? srt0 : srt1;
            const T R02 = R0 * R0, R06 = R02 * R02 * R02;
            const T r4 = r2 * r2, r6 = r4 * r2;
            const T t6 = r6 + R06, t8 = fma_t(r6, r2, R06 * R02);
            // one reciprocal for 1/t6 and 1/t8
            const T ip = rcp_pair(t6 * t8);
            const T it6 = t8 * ip, it8 = t6 * ip;
            const T phi = fma_t(ev, rinv, -(zp * (r - rc)));
            const T dispu = -fma_t(sr, it8, it6);
            const T zr = jo ? zrt0 : zrt1, al = jo ? alt0 : alt1;
            const T rep = zr * exp_pair(-(al * r)) * rinv;
            const T qq = qi * qj;
            e += 0.5 * (double)fma_t(qq, phi, fma_t(c6, dispu, rep));
            if (GRAD || PAR) {
                const T dphi = fma_t(-(ev * rinv), rinv, -fma_t(two_over_spi * za * ex2, rinv, zp));
                const T ddisp = fma_t(sr * (T)8 * r6 * r * it8, it8, (T)6 * r4 * r * it6 * it6);
                const T drep = -(rep * (al + rinv));
                if (GRAD) {
                    const T fp = fma_t(qq, dphi, fma_t(c6, ddisp, drep)) * rinv;
                    g0.add(fp * d0);
                    g1.add(fp * d1);
                    g2.add(fp * d2);
                }
                if (PAR) {
                    // d/dq_type of q_i q_j: own charge gets q_j, partner's charge gets q_i
                    const double hj = 0.5 * (double)(qj * phi), hi = 0.5 * (double)(qi * phi);
                    if (ti == 0) dq0 += hj; else dq1 += hj;
                    if (jo) dq0 += hi; else dq1 += hi;
                    const int cp = ti + tj;
                    const double hd = 0.5 * (double)dispu;
                    if (cp == 0) dc0 += hd; else if (cp == 1) dc1 += hd; else dc2 += hd;
                    if (vb) {
                        const T proj =
                            fma_t(vi0 - sv[jj][0], d0, fma_t(vi1 - sv[jj][1], d1, (vi2 - sv[jj][2]) * d2)) * rinv;
                        const double vj_ = 0.5 * (double)(qj * dphi * proj), vi_ = 0.5 * (double)(qi * dphi * proj);
                        if (ti == 0) dvq0 += vj_; else dvq1 += vj_;
                        if (jo) dvq0 += vi_; else dvq1 += vi_;
                        const double vd = 0.5 * (double)(ddisp * proj);
                        if (cp == 0) dvc0 += vd; else if (cp == 1) dvc1 += vd; else dvc2 += vd;
                    }
                }
            }
    };
    // this lane's part of its row for the split (LIST), or whole tiles
    bool scan = !LIST;
    int k = 0, kend = 0;
    const int *row = pl.row + (size_t)s * pl.rowcap;
    if (LIST && act) {
        const int *ss = pl.sst + (size_t)s * (S + 1);
        if (ss[0] < 0) scan = true;
        else {
            k = ss[split];
            kend = ss[split + 1];
        }
    }
    int jn = k < kend ? row[k] : 0x7FFFFFFF;
    const unsigned *mrow = pl.mask + (size_t)brick * pl.maskw;
    const int t0 = (int)((long long)split * pl.ntiles / S), t1 = (int)((long long)(split + 1) * pl.ntiles / S);
    for (int c0 = t0; c0 < t1; c0 += kChunkTiles) {
        const int c1 = min(t1, c0 + kChunkTiles);
        unsigned used = (1u << (c1 - c0)) - 1u;        // the same for the whole block
        if (LIST) {
            used = 0u;
            for (int t = c0; t < c1; ++t) used |= ((mrow[t >> 5] >> (t & 31)) & 1u) << (t - c0);
        }
        if (!used) continue;
        const int jb = c0 * kPTile, je = min(n, c1 * kPTile);
        __syncthreads();
        for (int q = tid; q < je - jb; q += kBrick) {
            if (!((used >> (q / kPTile)) & 1u)) continue;
            const int j = jb + q;
            const int jo = pl.perm[j];
            sx[q][0] = (T)xb[3 * jo];
            sx[q][1] = (T)xb[3 * jo + 1];
            sx[q][2] = (T)xb[3 * jo + 2];
            smt[q] = pl.smt[j];
            if (PAR) {
                sv[q][0] = vb ? (T)vb[3 * jo] : T(0);
                sv[q][1] = vb ? (T)vb[3 * jo + 1] : T(0);
                sv[q][2] = vb ? (T)vb[3 * jo + 2] : T(0);
            }
        }
        __syncthreads();
        if (!act) continue;
        if (scan) {
            for (int jj = 0; jj < je - jb; ++jj) {
                if ((smt[jj] >> 1) == mi) continue;     // same molecule (and the atom itself)
                T d0, d1, d2;
                const T r2 = sep(jj, d0, d1, d2);
                if (r2 <= rc2) pair(jj, d0, d1, d2, r2);
            }
        } else {
            while (jn < je) {                           // the next entry is already requested
                const int jj = jn - jb;
                ++k;
                jn = k < kend ? row[k] : 0x7FFFFFFF;
                T d0, d1, d2;
                const T r2 = sep(jj, d0, d1, d2);
                if (r2 <= rc2) pair(jj, d0, d1, d2, r2);
            }
        }
    }
    if (GRAD && act) {
        double *gp = gpart + (((size_t)split * gridDim.z + b) * n + i) * 3;
        gp[0] = g0.get();
        gp[1] = g1.get();
        gp[2] = g2.get();
    }
    // block sums of energy and parameter terms: butterflies within each wave,
    // then the four waves in order
    const int nblk = gridDim.x * gridDim.y;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    double *ep = epart + ((size_t)b * (nblk + 1) + blk) * kNE;
    constexpr int nv = PAR ? 11 : 1;
    const double vals[11] = { e, dq0, dq1, dc0, dc1, dc2, dvq0, dvq1, dvc0, dvc1, dvc2 };
    const int wave = tid >> 6;
#pragma unroll
    for (int q = 0; q < nv; ++q) {
        double w = act ? vals[q] : 0.0;
        for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
        if ((tid & 63) == 0) red[wave][q] = w;
    }
    __syncthreads();
    if (tid < kNE) {
        // slot -> value: 0 energy, 6..10 dq/dc, 16..20 dvq/dvc, the rest zero
        int q = -1;
        if (tid == 0) q = 0;
        else if (PAR && tid >= 6 && tid <= 10) q = tid - 5;
        else if (PAR && tid >= 1 + kNP + 5) q = tid - (1 + kNP + 5) + 6;
        if (q >= 0) ep[tid] = ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
        else if (PAR) ep[tid] = 0.0;
    }
}

// FP64 at the register count the compiler picks (four waves per SIMD);
// binary32 forces and energies held to six waves per SIMD.
template <typename T, bool GRAD, bool PAR, bool LIST, bool WIDE>
__global__ __launch_bounds__(kBrick) void nonbonded_kernel(FFDev f, PairDev pl, Theta th, const double *__restrict__ x,
                                                           const double *__restrict__ v, double *__restrict__ gpart,
                                                           double *__restrict__ epart)
{
    nonbonded_body<T, GRAD, PAR, LIST, WIDE>(f, pl, th, x, v, gpart, epart);
}

template <bool GRAD, bool LIST, bool WIDE>
__global__ __launch_bounds__(kBrick) __attribute__((amdgpu_waves_per_eu(6, 8))) void nonbonded_kernel_f32(
    FFDev f, PairDev pl, Theta th, const double *__restrict__ x, const double *__restrict__ v,
    double *__restrict__ gpart, double *__restrict__ epart)
{
    nonbonded_body<float, GRAD, false, LIST, WIDE>(f, pl, th, x, v, gpart, epart);
}

template <typename T, bool GRAD, bool PAR, bool LIST, bool WIDE>
void launch_nonbonded(dim3 grid, hipStream_t s, const FFDev &f, const PairDev &pl, const Theta &th, const double *x,
                      const double *v, double *gpart, double *epart)
{
    if constexpr (sizeof(T) == 4 && !PAR)
        hipLaunchKernelGGL((nonbonded_kernel_f32<GRAD, LIST, WIDE>), grid, dim3(kBrick), 0, s, f, pl, th, x, v, gpart,
                           epart);
    else
        hipLaunchKernelGGL((nonbonded_kernel<T, GRAD, PAR, LIST, WIDE>), grid, dim3(kBrick), 0, s, f, pl, th, x, v,
                           gpart, epart);
}

// Bonded terms, one thread per (atom, bead): the thread evaluates every bond
// and angle its atom takes part in (from a CSR list) and adds the atom's own
// share of the gradient to the summed nonbonded partials.  Energies and
// parameter derivatives of a term are booked by the thread of its first atom.
template <bool GRAD, bool PAR>
__global__ __launch_bounds__(kTile) void bonded_kernel(FFDev f, Theta th, const double *__restrict__ x,
                                                       const double *__restrict__ v, int nsplit,
                                                       const double *__restrict__ gpart, double *__restrict__ g,
                                                       double *__restrict__ epart, int nblk_nb)
{
    const int n = f.natoms;
    const int tid = threadIdx.x;
    const int a = blockIdx.x * kTile + tid;
    const int b = blockIdx.y;
    const bool act = a < n;
    const double *xb = x + (size_t)b * n * 3;
    const double *vb = v ? v + (size_t)b * n * 3 : nullptr;
    double gg[3] = { 0, 0, 0 }, e = 0.0, dth[kNP], dvth[kNP];
    for (int p = 0; p < kNP; ++p) dth[p] = dvth[p] = 0.0;
    if (act) {
        if (GRAD)
            for (int s = 0; s < nsplit; ++s) {
                const double *gp = gpart + (((size_t)s * gridDim.y + b) * n + a) * 3;
                gg[0] += gp[0];
                gg[1] += gp[1];
                gg[2] += gp[2];
            }
        const double De = th.p[0], aa = th.p[1], r0 = th.p[2], kk = th.p[3], t0 = th.p[4];
        for (int t = f.csr_ptr[a]; t < f.csr_ptr[a + 1]; ++t) {
            const int code = f.csr_term[t];
            const int kind = code >> 30, role = code & 3, idx = (code >> 2) & 0x0FFFFFFF;
            const bool book = role == 0;
            if (kind == 0) {
                const int i = f.bonds[2 * idx], j = f.bonds[2 * idx + 1];
                double d[3], r = 0.0;
                for (int c = 0; c < 3; ++c) {
                    d[c] = xb[3 * i + c] - xb[3 * j + c];
                    d[c] -= f.box * rint(d[c] / f.box);
                    r += d[c] * d[c];
                }
                r = sqrt(r);
                const double xx = r - r0, ex = exp(-aa * xx), om = 1.0 - ex;
                const double fp = 2.0 * De * aa * ex * om;
                const double sgn = role == 0 ? 1.0 : -1.0;
                if (GRAD)
                    for (int c = 0; c < 3; ++c) gg[c] += sgn * fp * d[c] / r;
                if (book) {
                    e += De * om * om;
                    if (PAR) {
                        dth[0] += om * om;
                        dth[1] += 2.0 * De * om * xx * ex;
                        dth[2] += -2.0 * De * om * aa * ex;
                        if (vb) {
                            double proj = 0.0;
                            for (int c = 0; c < 3; ++c) proj += (vb[3 * i + c] - vb[3 * j + c]) * d[c] / r;
                            dvth[0] += 2.0 * aa * ex * om * proj;
                            dvth[1] += 2.0 * De * ex * (om * (1.0 - aa * xx) + aa * xx * ex) * proj;
                            dvth[2] += 2.0 * De * aa * aa * ex * (1.0 - 2.0 * ex) * proj;
                        }
                    }
                }
            } else {
                const int i = f.angles[3 * idx], j = f.angles[3 * idx + 1], k = f.angles[3 * idx + 2];
                double u[3], w[3], ru = 0.0, rw = 0.0, dot = 0.0;
                for (int c = 0; c < 3; ++c) {
                    u[c] = xb[3 * i + c] - xb[3 * j + c];
                    u[c] -= f.box * rint(u[c] / f.box);
                    w[c] = xb[3 * k + c] - xb[3 * j + c];
                    w[c] -= f.box * rint(w[c] / f.box);
                    ru += u[c] * u[c];
                    rw += w[c] * w[c];
                    dot += u[c] * w[c];
                }
                ru = sqrt(ru);
                rw = sqrt(rw);
                const double cs = dot / (ru * rw);
                const double dt = acos(cs) - t0;
                const double de = 2.0 * kk * dt;
                const double s = -1.0 / sqrt(1.0 - cs * cs);
                double gi[3], gk[3];
                for (int c = 0; c < 3; ++c) {
                    gi[c] = s * (w[c] / rw - cs * u[c] / ru) / ru;
                    gk[c] = s * (u[c] / ru - cs * w[c] / rw) / rw;
                }
                if (GRAD)
                    for (int c = 0; c < 3; ++c)
                        gg[c] += role == 0 ? de * gi[c] : role == 2 ? de * gk[c] : -de * (gi[c] + gk[c]);
                if (book) {
                    e += kk * dt * dt;
                    if (PAR) {
                        dth[3] += dt * dt;
                        dth[4] += -2.0 * kk * dt;
                        if (vb) {
                            double proj = 0.0;
                            for (int c = 0; c < 3; ++c)
                                proj += vb[3 * i + c] * gi[c] + vb[3 * k + c] * gk[c] - vb[3 * j + c] * (gi[c] + gk[c]);
                            dvth[3] += 2.0 * dt * proj;
                            dvth[4] += -2.0 * kk * proj;
                        }
                    }
                }
            }
        }
        if (GRAD) {
            double *go = g + ((size_t)b * n + a) * 3;
            go[0] = gg[0];
            go[1] = gg[1];
            go[2] = gg[2];
        }
    }
    // block sums (the block is one wave: butterflies, a fixed order) into the
    // bonded rows, which follow every bead's nonbonded rows
    const int blk = blockIdx.x;
    double *bp = epart + ((size_t)gridDim.y * (nblk_nb + 1)) * kNE + ((size_t)b * gridDim.x + blk) * kNE;
    auto wave_sum = [&](double val) -> double {
        double w = act ? val : 0.0;
        for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
        return w;
    };
    const double s0 = wave_sum(e);
    if (tid == 0) bp[0] = s0;
    if (PAR) {
#pragma unroll
        for (int p = 0; p < kNP; ++p) {
            const double a = wave_sum(dth[p]), c = wave_sum(dvth[p]);
            if (tid == 0) {
                bp[1 + p] = a;
                bp[1 + kNP + p] = c;
            }
        }
    }
}

// Per-bead totals, one wave per (slot, bead): the nonbonded block rows then
// the bonded block rows, lane k taking rows k, k + 64, ..., then butterflies
// (a fixed order).
__global__ __launch_bounds__(64) void ff_reduce_kernel(const double *epart, int nblk_nb, int nblk_b, int nbeads,
                                                       double *e, double *dth, double *dvth)
{
    const int slot = blockIdx.x, b = blockIdx.y, lane = threadIdx.x;
    const double *nb = epart + (size_t)b * (nblk_nb + 1) * kNE;
    const double *bd = epart + ((size_t)nbeads * (nblk_nb + 1)) * kNE + (size_t)b * nblk_b * kNE;
    double s = 0.0;
    for (int k = lane; k < nblk_nb + nblk_b; k += 64)
        s += k < nblk_nb ? nb[(size_t)k * kNE + slot] : bd[(size_t)(k - nblk_nb) * kNE + slot];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (lane != 0) return;
    if (slot == 0) {
        if (e) e[b] = s;
    } else if (slot <= kNP) {
        if (dth) dth[(size_t)b * kNP + slot - 1] = s;
    } else {
        if (dvth) dvth[(size_t)b * kNP + slot - 1 - kNP] = s;
    }
}


// ---------------------------------------------------------------------------
// The reference's flexible water model (src/egrad_water.f90:36-333), atoms in
// the order O, H, H per molecule.  Intermolecular pairs: Coulomb, plain or
// the Zahn form, with the reference's gradient e0/r^2 * d (the derivative of
// q_i q_j / r); LJ whenever the lower-index atom of the pair is an oxygen
// (the reference tests name(i) twice, :293).  With periodic boxes: minimum
// image and the cutoff r < rc on both terms.
// ---------------------------------------------------------------------------
struct WaterDev {
    int natoms, periodic, zahn;
    double box, rc, za, zp;
    double r0, r0hh, de, a, khh, krt, krr, qh, qo, sig, eps;
};

template <typename T, bool GRAD>
__global__ __launch_bounds__(kTile) void water_nb_kernel(WaterDev w, const double *__restrict__ x, int nsplit,
                                                         double *__restrict__ gpart, double *__restrict__ epart)
{
    __shared__ T sx[kTile][3];
    __shared__ double red[kTile];
    const int n = w.natoms;
    const int tid = threadIdx.x;
    const int i = blockIdx.x * kTile + tid;
    const int split = blockIdx.y, b = blockIdx.z;
    const double *xb = x + (size_t)b * n * 3;
    const bool act = i < n;
    const int ii = act ? i : 0;
    const T xi0 = (T)xb[3 * ii], xi1 = (T)xb[3 * ii + 1], xi2 = (T)xb[3 * ii + 2];
    const int mi = ii / 3;
    const bool io = ii % 3 == 0;
    const T box = (T)w.box, rc = (T)w.rc, za = (T)w.za, zp = (T)w.zp;
    const T qi = (T)(io ? w.qo : w.qh), qo = (T)w.qo, qh = (T)w.qh;
    const T sig2 = (T)(w.sig * w.sig), eps = (T)w.eps;
    Acc<T> g0, g1, g2;
    double e = 0.0;
    const int chunk = ((n + nsplit - 1) / nsplit + kTile - 1) / kTile * kTile;
    const int j0 = split * chunk, j1 = min(n, j0 + chunk);
    for (int jt = j0; jt < j1; jt += kTile) {
        __syncthreads();
        const int j = jt + tid;
        if (j < j1) {
            sx[tid][0] = (T)xb[3 * j];
            sx[tid][1] = (T)xb[3 * j + 1];
            sx[tid][2] = (T)xb[3 * j + 2];
        }
        __syncthreads();
        if (!act) continue;
        const int nt = min(kTile, j1 - jt);
        for (int jj = 0; jj < nt; ++jj) {
            const int jg = jt + jj;
            if (jg / 3 == mi) continue;
            T d0 = xi0 - sx[jj][0], d1 = xi1 - sx[jj][1], d2 = xi2 - sx[jj][2];
            if (w.periodic) {
                d0 = min_image(d0, box);
                d1 = min_image(d1, box);
                d2 = min_image(d2, box);
            }
            const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
            const T r = sqrt_t(r2);
            if (w.periodic && !(r < rc)) continue;
            const T rinv = (T)1 / r;
            const T qj = jg % 3 == 0 ? qo : qh;
            const T e0 = w.zahn ? qi * qj * (erfc_t(za * r) * rinv - zp * (r - rc)) : qi * qj * rinv;
            T fp = e0 * rinv * rinv;                  // reference gradient of the Coulomb term
            T ep = e0;
            // LJ when the lower-index atom of the pair is an oxygen
            if ((i < jg) ? io : (jg % 3 == 0)) {
                const T s2 = sig2 * rinv * rinv;
                const T s6 = s2 * s2 * s2;
                ep += (T)4 * eps * (s6 * s6 - s6);
                fp += (T)24 * eps * rinv * rinv * s6 * ((T)2 * s6 - (T)1);
            }
            e += 0.5 * (double)ep;
            if (GRAD) {
                g0.add(-fp * d0);
                g1.add(-fp * d1);
                g2.add(-fp * d2);
            }
        }
    }
    if (GRAD && act) {
        double *gp = gpart + (((size_t)split * gridDim.z + b) * n + i) * 3;
        gp[0] = g0.get();
        gp[1] = g1.get();
        gp[2] = g2.get();
    }
    const int nblk = gridDim.x * gridDim.y;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    double *ep = epart + ((size_t)b * (nblk + 1) + blk) * kNE;
    __syncthreads();
    red[tid] = act ? e : 0.0;
    __syncthreads();
    for (int s = kTile / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) {
        ep[0] = red[0];
        for (int k = 1; k < kNE; ++k) ep[k] = 0.0;
    }
}

// Intramolecular terms, one thread per (atom, bead): the thread evaluates its
// molecule (three distances) and adds its atom's share to the summed
// nonbonded partials; the oxygen's thread books the molecule's energy.
template <bool GRAD>
__global__ __launch_bounds__(kTile) void water_intra_kernel(WaterDev w, const double *__restrict__ x, int nsplit,
                                                            const double *__restrict__ gpart, double *__restrict__ g,
                                                            double *__restrict__ epart, int nblk_nb)
{
    __shared__ double red[kTile];
    const int n = w.natoms;
    const int tid = threadIdx.x;
    const int a = blockIdx.x * kTile + tid;
    const int b = blockIdx.y;
    const bool act = a < n;
    const double *xb = x + (size_t)b * n * 3;
    double e = 0.0;
    if (act) {
        double gg[3] = { 0.0, 0.0, 0.0 };
        if (GRAD)
            for (int s = 0; s < nsplit; ++s) {
                const double *gp = gpart + (((size_t)s * gridDim.y + b) * n + a) * 3;
                gg[0] += gp[0];
                gg[1] += gp[1];
                gg[2] += gp[2];
            }
        const int o = a / 3 * 3, role = a % 3;
        double v1[3], v2[3], vh[3], r1 = 0.0, r2 = 0.0, rh = 0.0;
        for (int c = 0; c < 3; ++c) {
            v1[c] = xb[3 * o + c] - xb[3 * (o + 1) + c];
            v2[c] = xb[3 * o + c] - xb[3 * (o + 2) + c];
            vh[c] = xb[3 * (o + 1) + c] - xb[3 * (o + 2) + c];
            if (w.periodic) {
                v1[c] -= w.box * rint(v1[c] / w.box);
                v2[c] -= w.box * rint(v2[c] / w.box);
                vh[c] -= w.box * rint(vh[c] / w.box);
            }
            r1 += v1[c] * v1[c];
            r2 += v2[c] * v2[c];
            rh += vh[c] * vh[c];
        }
        r1 = sqrt(r1);
        r2 = sqrt(r2);
        rh = sqrt(rh);
        const double d1 = r1 - w.r0, d2 = r2 - w.r0, dh = rh - w.r0hh;
        const double e1 = exp(w.a * d1), e2 = exp(w.a * d2);
        if (role == 0)
            e = w.de * (1.0 - e1) * (1.0 - e1) + w.de * (1.0 - e2) * (1.0 - e2) + 0.5 * w.khh * dh * dh
                + w.krt * dh * (d1 + d2) + w.krr * d1 * d2;
        if (GRAD) {
            const double gr1 = -2.0 * w.a * w.de * e1 * (1.0 - e1) + w.krt * dh + w.krr * d2;   // dE/dr_OH1
            const double gr2 = -2.0 * w.a * w.de * e2 * (1.0 - e2) + w.krt * dh + w.krr * d1;
            const double grh = w.khh * dh + w.krt * (d1 + d2);
            for (int c = 0; c < 3; ++c) {
                const double a1 = gr1 * v1[c] / r1, a2 = gr2 * v2[c] / r2, ah = grh * vh[c] / rh;
                gg[c] += role == 0 ? a1 + a2 : role == 1 ? ah - a1 : -a2 - ah;
            }
            double *go = g + ((size_t)b * n + a) * 3;
            go[0] = gg[0];
            go[1] = gg[1];
            go[2] = gg[2];
        }
    }
    __syncthreads();
    red[tid] = act ? e : 0.0;
    __syncthreads();
    for (int s = kTile / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    double *bp = epart + ((size_t)gridDim.y * (nblk_nb + 1)) * kNE + ((size_t)b * gridDim.x + blockIdx.x) * kNE;
    if (tid == 0) {
        bp[0] = red[0];
        for (int k = 1; k < kNE; ++k) bp[k] = 0.0;
    }
}

}  // namespace
}  // namespace cam

using namespace cam;

struct cam_ff {
    int kind = 0;                     // 0: QMDFF form, 1: the reference's water model, 2: the reference's QMDFF
    QmdffDev *qm = nullptr;           // kind 2
    WaterDev w{};
    FFDev d{};
    int natoms = 0;
    int *type = nullptr, *mol = nullptr, *bonds = nullptr, *angles = nullptr, *csr_ptr = nullptr, *csr_term = nullptr;
    double *gpart = nullptr, *epart = nullptr;
    int cap_beads = 0, nsplit = 4;    // the water model's pair splits
    size_t gpart_bytes = 0, epart_bytes = 0;
    int pair_mode = CAM_FF_PAIRS_LIST;
    PairDev pl{};                     // kind 0: sorted order, rows and their reference positions
    int xref_beads = 0;
    std::vector<void *> pl_allocs;
};

namespace {

// The range of erfc's argument a r over the pairs inside the cutoff, a rc
// (held at 6 beyond, where erfc < 2.2e-17): it selects the pair kernel's
// erfcx polynomial, on [0, 2.5] or on [0, 6] (erfc_tables.hpp).
void erfc_range(FFDev &d)
{
    d.exmax = std::min(6.0, std::max(1e-3, d.za * d.rc));
}

// Splits of the pair kernel's tiles for nbeads beads: enough blocks to fill
// the GPU (about 2048 of 256 lanes), at most one per tile.
int pair_splits(const PairDev &p, int nbeads)
{
    const long long want = (2048 + (long long)p.nbricks * nbeads - 1) / ((long long)p.nbricks * nbeads);
    return (int)std::max(1LL, std::min<long long>(want, std::min(kMaxSplit, p.ntiles)));
}

// The sorted order, rows and cell tables of a box of the QMDFF form, sized
// once for its atom count (the rows for its density with a bead spread of up
// to 1.5 bohr; a longer row makes its atom walk whole tiles).
int pairs_create(cam_ff *ff)
{
    PairDev &p = ff->pl;
    const int n = ff->natoms;
    p.n = n;
    p.box = ff->d.box;
    p.rlist = ff->d.rc + kSkin + kMargin;
    p.ntiles = (n + kPTile - 1) / kPTile;
    p.nbricks = (n + kBrick - 1) / kBrick;
    p.maskw = (p.ntiles + 31) / 32;
    const double dens = n / (p.box * p.box * p.box), rl = p.rlist + 3.0;
    const double est = 1.3 * dens * 4.18879020478639 * rl * rl * rl + 64.0;
    p.rowcap = (int)std::min<double>(n, est);
    p.rowcap = std::max(4, (p.rowcap + 3) / 4 * 4);
    p.type = ff->type;
    p.mol = ff->mol;
    const int nchk = (n + 3) / 4, nc3 = kMaxCellsDim * kMaxCellsDim * kMaxCellsDim;
    auto alloc = [&](void **ptr, size_t bytes) -> int {
        CAM_HIP(hipMalloc(ptr, std::max<size_t>(bytes, 16)));
        ff->pl_allocs.push_back(*ptr);
        return CAM_OK;
    };
    int rc = CAM_OK;
#define CAM_PA(field, count, type) \
    if (rc == CAM_OK) rc = alloc((void **)&p.field, (size_t)(count) * sizeof(type));
    CAM_PA(perm, n, int) CAM_PA(scell, n, int) CAM_PA(smt, n, int) CAM_PA(tcell, n, int)
    CAM_PA(scen, 4 * (size_t)n, double) CAM_PA(cstart, nc3, int) CAM_PA(ccount, nc3, int)
    CAM_PA(row, (size_t)n * p.rowcap, int) CAM_PA(sst, (size_t)n * (kMaxSplit + 1), int)
    CAM_PA(mask, (size_t)p.nbricks * p.maskw, unsigned) CAM_PA(cen, 4 * (size_t)n, double)
    CAM_PA(bdev, nchk, double) CAM_PA(bover, nchk, int)
    CAM_PA(ctl, kCtlWords, int)
#undef CAM_PA
    if (rc != CAM_OK) return rc;
    CAM_HIP(hipMemset(p.ctl, 0, kCtlWords * sizeof(int)));
    return CAM_OK;
}

int ff_ensure(cam_ff *ff, int nbeads)
{
    if (ff->kind == 2) return CAM_OK;             // the QMDFF tables hold their own buffers
    const int n = ff->natoms;
    const int nbx = (n + kTile - 1) / kTile;
    size_t gbytes, ebytes;
    if (ff->kind == 0) {
        const int S = pair_splits(ff->pl, nbeads);
        gbytes = (size_t)S * nbeads * n * 3 * sizeof(double);
        ebytes = ((size_t)nbeads * (ff->pl.nbricks * S + 1) + (size_t)nbeads * nbx) * kNE * sizeof(double);
        if (nbeads > ff->xref_beads) {            // reference positions of the rows
            if (ff->pl.xref) CAM_HIP(hipFree(ff->pl.xref));
            ff->pl.xref = nullptr;
            CAM_HIP(hipMalloc(&ff->pl.xref, (size_t)nbeads * n * 3 * sizeof(double)));
            CAM_HIP(hipMemset(ff->pl.xref, 0, (size_t)nbeads * n * 3 * sizeof(double)));
            ff->xref_beads = nbeads;
        }
    } else {
        gbytes = (size_t)ff->nsplit * nbeads * n * 3 * sizeof(double);
        ebytes = ((size_t)nbeads * (nbx * ff->nsplit + 1) + (size_t)nbeads * nbx) * kNE * sizeof(double);
    }
    if (gbytes > ff->gpart_bytes) {
        if (ff->gpart) CAM_HIP(hipFree(ff->gpart));
        ff->gpart = nullptr;
        CAM_HIP(hipMalloc(&ff->gpart, gbytes));
        ff->gpart_bytes = gbytes;
    }
    if (ebytes > ff->epart_bytes) {
        if (ff->epart) CAM_HIP(hipFree(ff->epart));
        ff->epart = nullptr;
        CAM_HIP(hipMalloc(&ff->epart, ebytes));
        ff->epart_bytes = ebytes;
    }
    return CAM_OK;
}

template <typename T>
int ff_launch(cam_ff *ff, const Theta &th, const double *x, int nbeads, double *e, double *g, double *dth,
              const double *v, double *dvth, hipStream_t s)
{
    int rc = ff_ensure(ff, nbeads);
    if (rc != CAM_OK) return rc;
    const int n = ff->natoms;
    const int nbx = (n + kTile - 1) / kTile;
    const bool grad = g != nullptr;
    const bool par = dth != nullptr || dvth != nullptr;
    if (ff->kind == 2) return qmdff_launch<T>(ff->qm, x, nbeads, e, g, s);
    if (ff->kind == 1) {
        const int nblk_nb = nbx * ff->nsplit;
        dim3 gnb(nbx, ff->nsplit, nbeads);
        if (grad)
            hipLaunchKernelGGL((water_nb_kernel<T, true>), gnb, dim3(kTile), 0, s, ff->w, x, ff->nsplit, ff->gpart,
                               ff->epart);
        else
            hipLaunchKernelGGL((water_nb_kernel<T, false>), gnb, dim3(kTile), 0, s, ff->w, x, ff->nsplit, ff->gpart,
                               ff->epart);
        CAM_LAUNCH_CHECK();
        if (grad)
            hipLaunchKernelGGL((water_intra_kernel<true>), dim3(nbx, nbeads), dim3(kTile), 0, s, ff->w, x, ff->nsplit,
                               ff->gpart, g, ff->epart, nblk_nb);
        else
            hipLaunchKernelGGL((water_intra_kernel<false>), dim3(nbx, nbeads), dim3(kTile), 0, s, ff->w, x,
                               ff->nsplit, ff->gpart, g, ff->epart, nblk_nb);
        CAM_LAUNCH_CHECK();
        hipLaunchKernelGGL(ff_reduce_kernel, dim3(1, nbeads), dim3(64), 0, s, ff->epart, nblk_nb, nbx, nbeads, e, nullptr,
                           nullptr);
        CAM_LAUNCH_CHECK();
        return CAM_OK;
    }
    // the sorted order and rows (rebuilt on the device when a bead moved more
    // than skin/2 or the bead count changed), then the pairs
    PairDev pl = ff->pl;
    pl.nb = nbeads;
    pl.nsplit = pair_splits(pl, nbeads);
    const int nsplit = pl.nsplit, nblk_nb = pl.nbricks * nsplit;
    pl.cg = 1;
    while (pl.cg * 2 <= std::min(nbeads, 64)) pl.cg *= 2;
    const int nchk = (n + 256 / pl.cg - 1) / (256 / pl.cg);
    hipLaunchKernelGGL(pl_check_kernel, dim3(nchk), dim3(256), 0, s, pl, x);
    CAM_LAUNCH_CHECK();
    hipLaunchKernelGGL(pl_sort_kernel, dim3(1), dim3(kSortThreads), 0, s, pl, nchk);
    CAM_LAUNCH_CHECK();
    hipLaunchKernelGGL(pl_rows_kernel, dim3(std::min((n + 3) / 4, 512)), dim3(256), 0, s, pl, x);
    CAM_LAUNCH_CHECK();
    dim3 gnb(pl.nbricks, nsplit, nbeads);
    const bool list = ff->pair_mode == CAM_FF_PAIRS_LIST;
    const bool wide = ff->d.exmax > 2.5;
#define CAM_NB(G, P, L)                                                                                        \
    do {                                                                                                       \
        if (wide) launch_nonbonded<T, G, P, L, true>(gnb, s, ff->d, pl, th, x, P ? v : nullptr, ff->gpart, ff->epart); \
        else launch_nonbonded<T, G, P, L, false>(gnb, s, ff->d, pl, th, x, P ? v : nullptr, ff->gpart, ff->epart); \
    } while (0)
    if (list) {
        if (grad && par) CAM_NB(true, true, true);
        else if (grad) CAM_NB(true, false, true);
        else if (par) CAM_NB(false, true, true);
        else CAM_NB(false, false, true);
    } else {
        if (grad && par) CAM_NB(true, true, false);
        else if (grad) CAM_NB(true, false, false);
        else if (par) CAM_NB(false, true, false);
        else CAM_NB(false, false, false);
    }
#undef CAM_NB
    CAM_LAUNCH_CHECK();
    dim3 gb(nbx, nbeads);
    if (grad && par)
        hipLaunchKernelGGL((bonded_kernel<true, true>), gb, dim3(kTile), 0, s, ff->d, th, x, v, nsplit, ff->gpart, g,
                           ff->epart, nblk_nb);
    else if (grad)
        hipLaunchKernelGGL((bonded_kernel<true, false>), gb, dim3(kTile), 0, s, ff->d, th, x, nullptr, nsplit,
                           ff->gpart, g, ff->epart, nblk_nb);
    else if (par)
        hipLaunchKernelGGL((bonded_kernel<false, true>), gb, dim3(kTile), 0, s, ff->d, th, x, v, nsplit, ff->gpart,
                           g, ff->epart, nblk_nb);
    else
        hipLaunchKernelGGL((bonded_kernel<false, false>), gb, dim3(kTile), 0, s, ff->d, th, x, nullptr, nsplit,
                           ff->gpart, g, ff->epart, nblk_nb);
    CAM_LAUNCH_CHECK();
    hipLaunchKernelGGL(ff_reduce_kernel, dim3(par ? kNE : 1, nbeads), dim3(64), 0, s, ff->epart, nblk_nb, nbx, nbeads, e,
                       dth, dvth);
    CAM_LAUNCH_CHECK();
    return CAM_OK;
}

}  // namespace

extern "C" {

int cam_ff_create(const cam_ff_desc *desc, cam_ff **out)
{
    CAM_CHECK_ARG(desc && out, "null argument");
    *out = nullptr;
    const int n = desc->natoms;
    CAM_CHECK_ARG(n >= 1, "natoms must be >= 1");
    CAM_CHECK_ARG(desc->type && desc->mol, "type and mol are required");
    CAM_CHECK_ARG(desc->nbonds >= 0 && desc->nangles >= 0, "negative term count");
    CAM_CHECK_ARG(desc->nbonds == 0 || desc->bonds, "bonds missing");
    CAM_CHECK_ARG(desc->nangles == 0 || desc->angles, "angles missing");
    CAM_CHECK_ARG(desc->box > 0.0 && desc->rcut > 0.0 && desc->rcut <= 0.5 * desc->box,
                  "need 0 < rcut <= box/2 (minimum image)");
    for (int i = 0; i < n; ++i) CAM_CHECK_ARG(desc->type[i] == 0 || desc->type[i] == 1, "type[%d] must be 0 or 1", i);
    for (int i = 0; i < 2 * desc->nbonds; ++i)
        CAM_CHECK_ARG(desc->bonds[i] >= 0 && desc->bonds[i] < n, "bond atom out of range");
    for (int i = 0; i < 3 * desc->nangles; ++i)
        CAM_CHECK_ARG(desc->angles[i] >= 0 && desc->angles[i] < n, "angle atom out of range");
    clear_error();
    // per-atom term lists: code = kind << 30 | term << 2 | role
    std::vector<std::vector<int>> terms(n);
    for (int b = 0; b < desc->nbonds; ++b)
        for (int r = 0; r < 2; ++r) terms[desc->bonds[2 * b + r]].push_back((0 << 30) | (b << 2) | r);
    for (int a = 0; a < desc->nangles; ++a)
        for (int r = 0; r < 3; ++r) terms[desc->angles[3 * a + r]].push_back((1 << 30) | (a << 2) | r);
    std::vector<int> ptr(n + 1, 0), lst;
    for (int i = 0; i < n; ++i) {
        ptr[i + 1] = ptr[i] + (int)terms[i].size();
        lst.insert(lst.end(), terms[i].begin(), terms[i].end());
    }
    if (lst.empty()) lst.push_back(0);
    std::unique_ptr<cam_ff> ff(new cam_ff());
    ff->natoms = n;
    CAM_HIP(hipMalloc(&ff->type, n * sizeof(int)));
    CAM_HIP(hipMalloc(&ff->mol, n * sizeof(int)));
    CAM_HIP(hipMalloc(&ff->bonds, std::max(1, 2 * desc->nbonds) * sizeof(int)));
    CAM_HIP(hipMalloc(&ff->angles, std::max(1, 3 * desc->nangles) * sizeof(int)));
    CAM_HIP(hipMalloc(&ff->csr_ptr, (n + 1) * sizeof(int)));
    CAM_HIP(hipMalloc(&ff->csr_term, lst.size() * sizeof(int)));
    CAM_HIP(hipMemcpy(ff->type, desc->type, n * sizeof(int), hipMemcpyHostToDevice));
    CAM_HIP(hipMemcpy(ff->mol, desc->mol, n * sizeof(int), hipMemcpyHostToDevice));
    if (desc->nbonds) CAM_HIP(hipMemcpy(ff->bonds, desc->bonds, 2 * desc->nbonds * sizeof(int), hipMemcpyHostToDevice));
    if (desc->nangles)
        CAM_HIP(hipMemcpy(ff->angles, desc->angles, 3 * desc->nangles * sizeof(int), hipMemcpyHostToDevice));
    CAM_HIP(hipMemcpy(ff->csr_ptr, ptr.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice));
    CAM_HIP(hipMemcpy(ff->csr_term, lst.data(), lst.size() * sizeof(int), hipMemcpyHostToDevice));
    FFDev &d = ff->d;
    d.natoms = n;
    d.nbonds = desc->nbonds;
    d.nangles = desc->nangles;
    d.box = desc->box;
    d.rc = desc->rcut;
    d.za = desc->zahn_a;
    d.zp = desc->zahn_par;
    erfc_range(d);
    std::memcpy(d.sr42, desc->sr42, sizeof d.sr42);
    std::memcpy(d.r0bj, desc->r0bj, sizeof d.r0bj);
    std::memcpy(d.zrep, desc->zrep, sizeof d.zrep);
    std::memcpy(d.arep, desc->arep, sizeof d.arep);
    d.type = ff->type;
    d.mol = ff->mol;
    d.bonds = ff->bonds;
    d.angles = ff->angles;
    d.csr_ptr = ff->csr_ptr;
    d.csr_term = ff->csr_term;
    const int prc = pairs_create(ff.get());
    if (prc != CAM_OK) {
        cam_ff_destroy(ff.release());
        return prc;
    }
    *out = ff.release();
    return CAM_OK;
}

int cam_water_create(int nwater, int periodic, double box, int zahn, double rcut, cam_ff **out)
{
    CAM_CHECK_ARG(out, "null argument");
    *out = nullptr;
    CAM_CHECK_ARG(nwater >= 1 && nwater <= (1 << 24), "nwater out of range (%d)", nwater);
    CAM_CHECK_ARG(!periodic || (box > 0.0 && rcut > 0.0 && rcut <= 0.5 * box),
                  "periodic boxes need 0 < rcut <= box/2 (minimum image)");
    CAM_CHECK_ARG(!zahn || periodic, "the Zahn form needs a periodic box");
    clear_error();
    std::unique_ptr<cam_ff> ff(new cam_ff());
    ff->kind = 1;
    ff->natoms = 3 * nwater;
    WaterDev &w = ff->w;
    w.natoms = 3 * nwater;
    w.periodic = periodic ? 1 : 0;
    w.zahn = zahn ? 1 : 0;
    w.box = periodic ? box : 0.0;
    w.rc = periodic ? rcut : 0.0;
    // Zahn parameters as set_periodic sets them (src/set_periodic.f90:93-97)
    const double bohr = 0.52917721092, hartree = 627.5094743;
    w.za = 0.2 * bohr;
    const double acut = w.za * w.rc;
    w.zp = w.zahn ? std::erfc(acut) / (w.rc * w.rc) + 2.0 * w.za / 1.7724538509055160 * std::exp(-acut * acut) / w.rc
                  : 0.0;
    // water_init (src/water_init.f90:83-106); k_rr and eps_OO are default REAL there
    w.r0 = 1.0 / bohr;
    w.r0hh = 1.633 / bohr;
    w.de = 101.9188 / hartree;
    w.a = 2.567 * bohr;
    w.khh = 328.645606 / hartree * bohr * bohr;
    w.krt = -211.4672 / hartree * bohr * bohr;
    w.krr = (double)111.70765f / hartree * bohr * bohr;
    w.qh = 0.41;
    w.qo = -0.82;
    w.sig = 3.166 / bohr;
    w.eps = (double)0.1554f / hartree;
    *out = ff.release();
    return CAM_OK;
}

int cam_qmdff_create(const cam_qmdff_desc *d, cam_ff **out)
{
    CAM_CHECK_ARG(d && out, "null argument");
    *out = nullptr;
    const int n = d->natoms;
    CAM_CHECK_ARG(n >= 1 && n <= 65536, "natoms must be in 1..65536");
    CAM_CHECK_ARG(d->nbond >= 0 && d->nangl >= 0 && d->ntors >= 0 && d->nhb >= 0 && d->nnci >= 0,
                  "negative term count");
    CAM_CHECK_ARG(d->at && d->molnum && d->q && d->c6xy && d->r094 && d->sr42 && d->zab && d->r0ab && d->rad &&
                      d->scalehb,
                  "missing per-atom or pair tables");
    CAM_CHECK_ARG((d->nbond == 0 || (d->bond && d->vbond)) && (d->nangl == 0 || (d->angl && d->vangl)) &&
                      (d->ntors == 0 || (d->tors && d->vtors)) && (d->nhb == 0 || (d->hb && d->vhb)) &&
                      (d->nnci == 0 || d->nci),
                  "missing term list");
    for (int i = 0; i < n; ++i) CAM_CHECK_ARG(d->at[i] >= 1 && d->at[i] <= 94, "at[%d] must be in 1..94", i);
    auto atom_ok = [&](const int32_t *v, size_t cnt) {
        for (size_t k = 0; k < cnt; ++k)
            if (v[k] < 1 || v[k] > n) return false;
        return true;
    };
    CAM_CHECK_ARG(atom_ok(d->bond, 2 * (size_t)d->nbond) && atom_ok(d->angl, 3 * (size_t)d->nangl) &&
                      atom_ok(d->hb, 3 * (size_t)d->nhb),
                  "atom index out of 1..natoms");
    for (int m = 0; m < d->ntors; ++m) {
        CAM_CHECK_ARG(atom_ok(d->tors + 6 * m, 4), "torsion %d: atom index out of range", m + 1);
        CAM_CHECK_ARG(d->tors[6 * m + 4] >= 0 && d->tors[6 * m + 4] <= 4, "torsion %d: 0..4 terms", m + 1);
    }
    for (int m = 0; m < d->nnci; ++m) {
        CAM_CHECK_ARG(atom_ok(d->nci + 3 * m, 2), "nci pair %d: atom index out of range", m + 1);
        CAM_CHECK_ARG(d->nci[3 * m + 2] >= 1 && d->nci[3 * m + 2] <= 6, "nci pair %d: kind must be 1..6", m + 1);
    }
    // halogen bonds (eabxag) are not implemented
    int nmols = 0;
    for (int i = 0; i < n; ++i) nmols = std::max(nmols, (int)d->molnum[i]);
    for (int m = 0; m < d->nhb; ++m)
        if (d->at[d->hb[3 * m + 2] - 1] != 1) {
            set_error("hydrogen-bond term %d has a halogen bridge (halogen bonds are not implemented)", m + 1);
            return CAM_ERR_UNSUPPORTED;
        }
    auto halogen = [](int z) { return z == 17 || z == 35 || z == 53 || z == 85; };
    if (nmols > 1)
        for (int m = 0; m < d->nbond; ++m) {
            const int a = d->at[d->bond[2 * m] - 1], b = d->at[d->bond[2 * m + 1] - 1];
            if ((halogen(a) && b != 1) || (halogen(b) && a != 1)) {
                set_error("bond %d gives intermolecular halogen bonds (not implemented)", m + 1);
                return CAM_ERR_UNSUPPORTED;
            }
        }
    if (d->periodic) {
        CAM_CHECK_ARG(d->box[0] > 0.0 && d->box[1] > 0.0 && d->box[2] > 0.0, "box lengths must be > 0");
    }
    clear_error();
    QmdffDev *q = nullptr;
    const int rc = qmdff_build(d, &q);
    if (rc != CAM_OK) return rc;
    cam_ff *ff = new cam_ff();
    ff->kind = 2;
    ff->qm = q;
    ff->natoms = n;
    *out = ff;
    return CAM_OK;
}

int cam_ff_destroy(cam_ff *ff)
{
    if (!ff) return CAM_OK;
    qmdff_free(ff->qm);
    int *ib[] = { ff->type, ff->mol, ff->bonds, ff->angles, ff->csr_ptr, ff->csr_term };
    for (int *p : ib)
        if (p) (void)hipFree(p);
    if (ff->gpart) (void)hipFree(ff->gpart);
    if (ff->epart) (void)hipFree(ff->epart);
    for (void *p : ff->pl_allocs) (void)hipFree(p);
    if (ff->pl.xref) (void)hipFree(ff->pl.xref);
    delete ff;
    return CAM_OK;
}

int cam_ff_set_pair_mode(cam_ff *ff, int mode)
{
    CAM_CHECK_ARG(ff, "null force field");
    CAM_CHECK_ARG(mode == CAM_FF_PAIRS_LIST || mode == CAM_FF_PAIRS_ALL, "bad pair mode %d", mode);
    clear_error();
    ff->pair_mode = mode;
    return CAM_OK;
}

int cam_ff_energy(cam_ff *ff, const double *theta, const double *x, int nbeads, double *e, int precision,
                  void *stream)
{
    CAM_CHECK_ARG(ff && x && e, "null argument");
    CAM_CHECK_ARG(theta || ff->kind != 0, "theta is required for this force field");
    CAM_CHECK_ARG(nbeads >= 1, "nbeads must be >= 1");
    CAM_CHECK_ARG(precision == CAM_PREC_FP64 || precision == CAM_PREC_MIXED, "bad precision");
    clear_error();
    Theta th{};
    if (theta) std::memcpy(th.p, theta, sizeof th.p);
    hipStream_t s = (hipStream_t)stream;
    return precision == CAM_PREC_MIXED ? ff_launch<float>(ff, th, x, nbeads, e, nullptr, nullptr, nullptr, nullptr, s)
                                       : ff_launch<double>(ff, th, x, nbeads, e, nullptr, nullptr, nullptr, nullptr, s);
}

int cam_ff_forces(cam_ff *ff, const double *theta, const double *x, int nbeads, double *e, double *dedx,
                  double *dedtheta, const double *v, double *dvdtheta, int precision, void *stream)
{
    CAM_CHECK_ARG(ff && x && dedx, "null argument");
    CAM_CHECK_ARG(theta || ff->kind != 0, "theta is required for this force field");
    CAM_CHECK_ARG(ff->kind == 0 || (!dedtheta && !dvdtheta), "this force field has no parameter derivatives");
    CAM_CHECK_ARG(nbeads >= 1, "nbeads must be >= 1");
    CAM_CHECK_ARG(!dvdtheta || v, "dvdtheta needs the direction v");
    CAM_CHECK_ARG(precision == CAM_PREC_FP64 || precision == CAM_PREC_MIXED, "bad precision");
    clear_error();
    Theta th{};
    if (theta) std::memcpy(th.p, theta, sizeof th.p);
    hipStream_t s = (hipStream_t)stream;
    return precision == CAM_PREC_MIXED
               ? ff_launch<float>(ff, th, x, nbeads, e, dedx, dedtheta, dvdtheta ? v : nullptr, dvdtheta, s)
               : ff_launch<double>(ff, th, x, nbeads, e, dedx, dedtheta, dvdtheta ? v : nullptr, dvdtheta, s);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Ring-polymer engine for one periodic box: nbeads replicas of all atoms.
// Per step: half kick + normal-mode step (box_nm_kernel), forces of every
// bead (the adjoint kernels above), half kick + Andersen (box_kick_kernel).
// ---------------------------------------------------------------------------
namespace cam {
namespace {

struct BoxNM {
    double *q, *p;
    const double *g;
    const double *mass;        // per atom
    const double *tfwd, *tinv; // [nb][nb], element (j,k) = T[k][j]
    int natoms, nb;
    double dt, beta;
};

// One block: 256/nb columns (atom, dim) x nb beads.
__global__ __launch_bounds__(256) void box_nm_kernel(BoxNM a)
{
    extern __shared__ double smem[];
    const int nb = a.nb, cpb = 256 / nb;
    const int ncol = 3 * a.natoms;
    const int tid = threadIdx.x;
    const int lc = tid / nb, k = tid % nb;
    const int col = blockIdx.x * cpb + lc;
    const bool act = col < ncol;
    double *sp = smem, *sq = smem + 256;
    const size_t o = (size_t)k * ncol + col;
    double p = 0.0, q = 0.0;
    if (act) {
        p = a.p[o] - 0.5 * a.dt * a.g[o];
        q = a.q[o];
    }
    sp[lc * nb + k] = p;
    sq[lc * nb + k] = q;
    __syncthreads();
    double yp = 0.0, yq = 0.0;
    for (int j = 0; j < nb; ++j) {
        const double t = a.tfwd[j * nb + k];
        yp = fma(t, sp[lc * nb + j], yp);
        yq = fma(t, sq[lc * nb + j], yq);
    }
    // exact free ring-polymer propagator of mode k (src/verlet.f90:405-432)
    const double m = act ? a.mass[col / 3] : 1.0;
    double pn, qn;
    if (k == 0 || nb == 1) {
        pn = yp;
        qn = yp * (a.dt / m) + yq;
    } else {
        const int kk = k <= nb / 2 ? k : nb - k;
        const double wk = (2.0 / (a.beta / nb)) * sin(kk * (3.141592653589793 / nb));
        const double wt = wk * a.dt, wm = wk * m;
        const double cw = cos(wt), sw = sin(wt);
        pn = yp * cw - yq * wm * sw;
        qn = yp * (sw / wm) + yq * cw;
    }
    __syncthreads();
    sp[lc * nb + k] = pn;
    sq[lc * nb + k] = qn;
    __syncthreads();
    double xp = 0.0, xq = 0.0;
    for (int j = 0; j < nb; ++j) {
        const double t = a.tinv[j * nb + k];
        xp = fma(t, sp[lc * nb + j], xp);
        xq = fma(t, sq[lc * nb + j], xq);
    }
    if (act) {
        a.p[o] = xp;
        a.q[o] = xq;
    }
}

__global__ void box_kick_kernel(double *p, const double *g, const double *mass, int natoms, int nb, double dt,
                                double beta, int andersen_step, uint64_t seed, const uint64_t *step_base, int step_off)
{
    const size_t n = (size_t)natoms * 3 * nb;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = p[i] - 0.5 * dt * g[i];
    if (andersen_step > 0) {
        const uint64_t istep = *step_base + (uint64_t)step_off + 1;
        if (istep % (uint64_t)andersen_step == 0) {
            const int k = (int)(i / ((size_t)natoms * 3));
            const int c = (int)(i % ((size_t)natoms * 3));
            v = andersen_normal(seed, 0u, istep, c, k) * sqrt(mass[c / 3] / (beta / nb));
        }
    }
    p[i] = v;
}

__global__ void box_advance_kernel(uint64_t *ctr, int n) { *ctr += (uint64_t)n; }

__global__ void box_energy_kernel(const double *q, const double *p, const double *mass, int natoms, int nb,
                                  double beta, double *ek, double *es)
{
    __shared__ double r1[256], r2[256];
    const size_t ncol = (size_t)natoms * 3;
    double a = 0.0, b = 0.0;
    const double wn = nb / beta;
    for (size_t i = threadIdx.x; i < ncol * nb; i += 256) {
        const int k = (int)(i / ncol);
        const size_t c = i % ncol;
        const double m = mass[c / 3];
        a += p[i] * p[i] / (2.0 * m);
        if (nb > 1) {
            const double dq = q[i] - q[(size_t)((k + 1) % nb) * ncol + c];
            b += 0.5 * m * wn * wn * dq * dq;
        }
    }
    r1[threadIdx.x] = a;
    r2[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r1[threadIdx.x] += r1[threadIdx.x + w];
            r2[threadIdx.x] += r2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        *ek = r1[0];
        *es = r2[0];
    }
}

// Nose-Hoover chain of the box (thermostat 2, src/nhc.f90:35-169): the
// centroid-momenta kinetic energy sum p^2/(2m)/nbeads^2 in per-block parts
// (a fixed order, so runs repeat bit for bit), the chain in one thread, then
// every momentum scaled.  st[0..3]: thermostat velocities, st[4]: the scale.
constexpr int kBoxNhcBlocks = 256;

__global__ __launch_bounds__(256) void box_ke_kernel(const double *p, const double *mass, int natoms, int nb,
                                                     double *part)
{
    __shared__ double red[256];
    const size_t n = (size_t)3 * natoms * nb;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double v = p[i];
        acc += v * v / (2.0 * mass[(i % (3 * (size_t)natoms)) / 3]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(64) void box_nhc_kernel(const double *part, int nparts, double *st, double beta,
                                                     double nose_q, int natoms, int nb, double dt)
{
    // the kinetic-energy parts: lane k takes parts k, k + 64, ..., then
    // butterflies (a fixed order); lane 0 runs the chain
    const int lane = threadIdx.x;
    double ek = 0.0;
    for (int i = lane; i < nparts; i += 64) ek += part[i];
    for (int o = 32; o > 0; o >>= 1) ek += __shfl_xor(ek, o);
    if (lane != 0 || blockIdx.x != 0) return;
    const double kelvin = 1.0 / (beta * 0.316679e-5);
    const double qterm = 0.316679e-5 * kelvin * nose_q * nose_q;    // mdinit (src/mdinit.f90:139-146)
    const int nfree = 3 * natoms;
    const double qnh[4] = { (double)nfree * qterm, qterm, qterm, qterm };
    const double ekt = (double)(1.380649E-23f / 4.3597447E-18f) * kelvin;
    double vnh[4] = { st[0], st[1], st[2], st[3] };
    const double scale = nhc_chain(vnh, qnh, ek / nb / nb, ekt, nfree, dt);
    for (int j = 0; j < 4; ++j) st[j] = vnh[j];
    st[4] = scale;
}

__global__ void box_scale_kernel(double *p, size_t n, const double *st)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = st[4] * p[i];
}

// transrot for one box (src/transrot.f90:36-237, with its total mass of
// nbeads^2 * M): two passes of fixed-order block sums over every (bead,
// atom) -- mass-weighted velocities and positions, then, about the centre
// they give, the angular momentum and the inertia tensor -- one thread that
// inverts the tensor, and the velocity update.  tr[0..2] vtot, [3..5] centre,
// [6..8] angular velocity.
constexpr int kTrBlocks = 64;

__device__ void tr_block_sum(double *v, int nv, double *sh, double *out)
{
    // v: this thread's nv partial sums -> out[nv] for the block, fixed order
    for (int c = 0; c < nv; ++c) sh[c * 256 + threadIdx.x] = v[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < nv; ++c) sh[c * 256 + threadIdx.x] = sh[c * 256 + threadIdx.x] + sh[c * 256 + threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < nv; ++c) out[c] = sh[c * 256];
}

__global__ __launch_bounds__(256) void box_tr_pass1(const double *q, const double *p, const double *mass, int natoms,
                                                    int nb, double *part)
{
    __shared__ double sh[6 * 256];
    double v[6] = { 0, 0, 0, 0, 0, 0 };
    const size_t n = (size_t)natoms * nb;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)kTrBlocks * 256) {
        const double w = mass[e % natoms];
        for (int c = 0; c < 3; ++c) {
            v[c] = v[c] + (p[3 * e + c] / w) * w;
            v[3 + c] = v[3 + c] + q[3 * e + c] * w;
        }
    }
    tr_block_sum(v, 6, sh, part + 6 * blockIdx.x);
}

__global__ __launch_bounds__(256) void box_tr_pass2(const double *q, const double *p, const double *mass, int natoms,
                                                    int nb, const double *part1, double totmass, double *tr,
                                                    double *part)
{
    __shared__ double sh[9 * 256];
    __shared__ double ctr[6];
    if (threadIdx.x == 0) {
        double s[6] = { 0, 0, 0, 0, 0, 0 };
        for (int b = 0; b < kTrBlocks; ++b)
            for (int c = 0; c < 6; ++c) s[c] = s[c] + part1[6 * b + c];
        for (int c = 0; c < 6; ++c) ctr[c] = s[c] / totmass;
        if (blockIdx.x == 0)
            for (int c = 0; c < 6; ++c) tr[c] = ctr[c];
    }
    __syncthreads();
    double v[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
    const size_t n = (size_t)natoms * nb;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)kTrBlocks * 256) {
        const double w = mass[e % natoms];
        const double *x = q + 3 * e;
        const double vx = p[3 * e] / w, vy = p[3 * e + 1] / w, vz = p[3 * e + 2] / w;
        v[0] = v[0] + (x[1] * vz - x[2] * vy) * w;
        v[1] = v[1] + (x[2] * vx - x[0] * vz) * w;
        v[2] = v[2] + (x[0] * vy - x[1] * vx) * w;
        const double xd = x[0] - ctr[3], yd = x[1] - ctr[4], zd = x[2] - ctr[5];
        v[3] = v[3] + xd * xd * w;
        v[4] = v[4] + xd * yd * w;
        v[5] = v[5] + xd * zd * w;
        v[6] = v[6] + yd * yd * w;
        v[7] = v[7] + yd * zd * w;
        v[8] = v[8] + zd * zd * w;
    }
    tr_block_sum(v, 9, sh, part + 9 * blockIdx.x);
}

__global__ void box_tr_solve(const double *part, int natoms, double totmass, double *tr)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
    for (int b = 0; b < kTrBlocks; ++b)
        for (int c = 0; c < 9; ++c) s[c] = s[c] + part[9 * b + c];
    const double vt[3] = { tr[0], tr[1], tr[2] }, xt = tr[3], yt = tr[4], zt = tr[5];
    double mang[3];
    mang[0] = s[0] - (yt * vt[2] - zt * vt[1]) * totmass;
    mang[1] = s[1] - (zt * vt[0] - xt * vt[2]) * totmass;
    mang[2] = s[2] - (xt * vt[1] - yt * vt[0]) * totmass;
    const double xx = s[3], xy = s[4], xz = s[5], yy = s[6], yz = s[7], zz = s[8];
    double ten[3][3];
    ten[0][0] = yy + zz; ten[0][1] = -xy;    ten[0][2] = -xz;
    ten[1][0] = -xy;     ten[1][1] = xx + zz; ten[1][2] = -yz;
    ten[2][0] = -xz;     ten[2][1] = -yz;    ten[2][2] = xx + yy;
    if (natoms <= 2) {
        ten[0][0] += 0.000001;
        ten[1][1] += 0.000001;
        ten[2][2] += 0.000001;
    }
    // Gauss-Jordan inversion with full pivoting (src/invert.f90)
    int ipiv[3] = { 0, 0, 0 }, indxr[3], indxc[3];
    for (int i = 0; i < 3; ++i) {
        double big = 0.0;
        int irow = 0, icol = 0;
        for (int j = 0; j < 3; ++j)
            if (ipiv[j] != 1)
                for (int kk = 0; kk < 3; ++kk)
                    if (ipiv[kk] == 0 && fabs(ten[j][kk]) >= big) {
                        big = fabs(ten[j][kk]);
                        irow = j;
                        icol = kk;
                    }
        ipiv[icol] += 1;
        if (irow != icol)
            for (int j = 0; j < 3; ++j) {
                const double tmp = ten[irow][j];
                ten[irow][j] = ten[icol][j];
                ten[icol][j] = tmp;
            }
        indxr[i] = irow;
        indxc[i] = icol;
        const double piv = ten[icol][icol];
        ten[icol][icol] = 1.0;
        for (int j = 0; j < 3; ++j) ten[icol][j] = ten[icol][j] / piv;
        for (int j = 0; j < 3; ++j)
            if (j != icol) {
                const double tmp = ten[j][icol];
                ten[j][icol] = 0.0;
                for (int kk = 0; kk < 3; ++kk) ten[j][kk] = ten[j][kk] - ten[icol][kk] * tmp;
            }
    }
    for (int i = 2; i >= 0; --i)
        if (indxr[i] != indxc[i])
            for (int kk = 0; kk < 3; ++kk) {
                const double tmp = ten[kk][indxr[i]];
                ten[kk][indxr[i]] = ten[kk][indxc[i]];
                ten[kk][indxc[i]] = tmp;
            }
    for (int i = 0; i < 3; ++i) {
        double v = 0.0;
        for (int j = 0; j < 3; ++j) v = v + ten[i][j] * mang[j];
        tr[6 + i] = v;
    }
}

__global__ void box_tr_apply(const double *q, double *p, const double *mass, int natoms, int nb, const double *tr)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)natoms * nb) return;
    const double w = mass[e % natoms];
    double v[3];
    for (int j = 0; j < 3; ++j) v[j] = p[3 * e + j] / w - tr[j];
    const double xd = q[3 * e] - tr[3], yd = q[3 * e + 1] - tr[4], zd = q[3 * e + 2] - tr[5];
    v[0] = v[0] - tr[7] * zd + tr[8] * yd;
    v[1] = v[1] - tr[8] * xd + tr[6] * zd;
    v[2] = v[2] - tr[6] * yd + tr[7] * xd;
    for (int j = 0; j < 3; ++j) p[3 * e + j] = v[j] * w;
}

}  // namespace
}  // namespace cam

struct cam_box {
    cam_ff *ff = nullptr;
    int natoms = 0, nb = 0, precision = CAM_PREC_FP64, andersen_step = 0;
    double dt = 0.0, beta = 1.0;
    uint64_t seed = 0;
    double theta[10] = {};
    hipStream_t stream = nullptr;
    double *q = nullptr, *p = nullptr, *g = nullptr, *e = nullptr, *mass = nullptr, *tfwd = nullptr, *tinv = nullptr,
           *ek = nullptr;
    uint64_t *step_base = nullptr;
    double nose_q = 0.0;               // Nose-Hoover chain on when > 0
    double *nhc = nullptr;             // chain state: 4 velocities and the last scale factor
    double *nhc_part = nullptr;        // kinetic-energy parts (kBoxNhcBlocks)
    bool transrot = false;             // remove centre-of-mass motion after every step (dynamic.x)
    double totmass = 0.0;              // transrot's nbeads^2 * M
    double *tr = nullptr;              // transrot: partial sums and the solved velocities
    bool have_state = false;
    std::map<int, hipGraphExec_t> graphs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evf0 = nullptr, evf1 = nullptr;
    double last_ms = -1.0;
};

namespace cam {
void nm_matrices(int nb, int mode, std::vector<double> &fwd, std::vector<double> &inv);
}

namespace {

int box_enqueue(cam_box *bx, int nsteps)
{
    const int ncol = 3 * bx->natoms;
    const int cpb = 256 / bx->nb;
    BoxNM a{ bx->q, bx->p, bx->g, bx->mass, bx->tfwd, bx->tinv, bx->natoms, bx->nb, bx->dt, bx->beta };
    Theta th;
    std::memcpy(th.p, bx->theta, sizeof th.p);
    const size_t n = (size_t)ncol * bx->nb;
    const unsigned egrid = (unsigned)((n + 255) / 256);
    auto chain = [&]() -> int {                        // src/verlet.f90:185-206, :1086-1106
        if (bx->nose_q <= 0.0) return CAM_OK;
        hipLaunchKernelGGL(box_ke_kernel, dim3(kBoxNhcBlocks), dim3(256), 0, bx->stream, bx->p, bx->mass, bx->natoms,
                           bx->nb, bx->nhc_part);
        CAM_LAUNCH_CHECK();
        hipLaunchKernelGGL(box_nhc_kernel, dim3(1), dim3(64), 0, bx->stream, bx->nhc_part, kBoxNhcBlocks, bx->nhc,
                           bx->beta, bx->nose_q, bx->natoms, bx->nb, bx->dt);
        CAM_LAUNCH_CHECK();
        hipLaunchKernelGGL(box_scale_kernel, dim3(egrid), dim3(256), 0, bx->stream, bx->p, n, bx->nhc);
        CAM_LAUNCH_CHECK();
        return CAM_OK;
    };
    for (int s = 0; s < nsteps; ++s) {
        int rc0 = chain();
        if (rc0 != CAM_OK) return rc0;
        hipLaunchKernelGGL(box_nm_kernel, dim3((ncol + cpb - 1) / cpb), dim3(256), 2 * 256 * sizeof(double), bx->stream,
                           a);
        CAM_LAUNCH_CHECK();
        if (bx->ff->kind == 2) {                       // the periodic wrap of src/verlet.f90:540-640
            const int rw = qmdff_wrap(bx->ff->qm, bx->q, bx->nb, bx->stream);
            if (rw != CAM_OK) return rw;
        }
        int rc = bx->precision == CAM_PREC_MIXED
                     ? ff_launch<float>(bx->ff, th, bx->q, bx->nb, bx->e, bx->g, nullptr, nullptr, nullptr, bx->stream)
                     : ff_launch<double>(bx->ff, th, bx->q, bx->nb, bx->e, bx->g, nullptr, nullptr, nullptr, bx->stream);
        if (rc != CAM_OK) return rc;
        hipLaunchKernelGGL(box_kick_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, bx->stream, bx->p, bx->g,
                           bx->mass, bx->natoms, bx->nb, bx->dt, bx->beta, bx->andersen_step, bx->seed, bx->step_base, s);
        CAM_LAUNCH_CHECK();
        rc = chain();
        if (rc != CAM_OK) return rc;
        if (bx->transrot) {                             // src/verlet.f90:1299-1305
            double *part1 = bx->tr + 16, *part2 = part1 + 6 * kTrBlocks;
            hipLaunchKernelGGL(box_tr_pass1, dim3(kTrBlocks), dim3(256), 0, bx->stream, bx->q, bx->p, bx->mass,
                               bx->natoms, bx->nb, part1);
            CAM_LAUNCH_CHECK();
            hipLaunchKernelGGL(box_tr_pass2, dim3(kTrBlocks), dim3(256), 0, bx->stream, bx->q, bx->p, bx->mass,
                               bx->natoms, bx->nb, part1, bx->totmass, bx->tr, part2);
            CAM_LAUNCH_CHECK();
            hipLaunchKernelGGL(box_tr_solve, dim3(1), dim3(64), 0, bx->stream, part2, bx->natoms, bx->totmass, bx->tr);
            CAM_LAUNCH_CHECK();
            const size_t ne = (size_t)bx->natoms * bx->nb;
            hipLaunchKernelGGL(box_tr_apply, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, bx->stream, bx->q,
                               bx->p, bx->mass, bx->natoms, bx->nb, bx->tr);
            CAM_LAUNCH_CHECK();
        }
    }
    hipLaunchKernelGGL(box_advance_kernel, dim3(1), dim3(1), 0, bx->stream, bx->step_base, nsteps);
    CAM_LAUNCH_CHECK();
    return CAM_OK;
}

}  // namespace

extern "C" {

int cam_box_create(cam_ff *ff, int nbeads, double dt, const double *mass, double beta, int nm_mode, int precision,
                   int andersen_step, uint64_t seed, cam_box **out)
{
    CAM_CHECK_ARG(ff && mass && out, "null argument");
    *out = nullptr;
    CAM_CHECK_ARG(nbeads >= 1 && nbeads <= 256 && (nbeads & (nbeads - 1)) == 0, "nbeads must be a power of two <= 256");
    CAM_CHECK_ARG(dt > 0.0 && beta > 0.0, "dt and beta must be > 0");
    CAM_CHECK_ARG(nm_mode == CAM_NM_REFERENCE || nm_mode == CAM_NM_ORTHOGONAL, "bad nm_mode");
    CAM_CHECK_ARG(precision == CAM_PREC_FP64 || precision == CAM_PREC_MIXED, "bad precision");
    CAM_CHECK_ARG(andersen_step >= 0, "andersen_step must be >= 0");
    for (int i = 0; i < ff->natoms; ++i) CAM_CHECK_ARG(mass[i] > 0.0, "mass[%d] must be > 0", i);
    clear_error();
    std::unique_ptr<cam_box> bx(new cam_box());
    bx->ff = ff;
    bx->natoms = ff->natoms;
    bx->nb = nbeads;
    bx->dt = dt;
    bx->beta = beta;
    bx->precision = precision;
    bx->andersen_step = andersen_step;
    bx->seed = seed;
    const size_t n = (size_t)bx->natoms * 3 * nbeads;
    CAM_HIP(hipStreamCreateWithFlags(&bx->stream, hipStreamNonBlocking));
    CAM_HIP(hipMalloc(&bx->q, n * sizeof(double)));
    CAM_HIP(hipMalloc(&bx->p, n * sizeof(double)));
    CAM_HIP(hipMalloc(&bx->g, n * sizeof(double)));
    CAM_HIP(hipMalloc(&bx->e, nbeads * sizeof(double)));
    CAM_HIP(hipMalloc(&bx->ek, 2 * sizeof(double)));
    CAM_HIP(hipMalloc(&bx->mass, bx->natoms * sizeof(double)));
    CAM_HIP(hipMalloc(&bx->step_base, sizeof(uint64_t)));
    CAM_HIP(hipMalloc(&bx->tfwd, (size_t)nbeads * nbeads * sizeof(double)));
    CAM_HIP(hipMalloc(&bx->tinv, (size_t)nbeads * nbeads * sizeof(double)));
    std::vector<double> fwd, inv;
    nm_matrices(nbeads, nm_mode, fwd, inv);
    CAM_HIP(hipMemcpy(bx->tfwd, fwd.data(), fwd.size() * sizeof(double), hipMemcpyHostToDevice));
    CAM_HIP(hipMemcpy(bx->tinv, inv.data(), inv.size() * sizeof(double), hipMemcpyHostToDevice));
    CAM_HIP(hipMemcpy(bx->mass, mass, bx->natoms * sizeof(double), hipMemcpyHostToDevice));
    CAM_HIP(hipMemset(bx->step_base, 0, sizeof(uint64_t)));
    CAM_HIP(hipEventCreate(&bx->ev0));
    CAM_HIP(hipEventCreate(&bx->ev1));
    *out = bx.release();
    return CAM_OK;
}

int cam_box_destroy(cam_box *bx)
{
    if (!bx) return CAM_OK;
    (void)hipStreamSynchronize(bx->stream);
    for (auto &kv : bx->graphs) (void)hipGraphExecDestroy(kv.second);
    double *bufs[] = { bx->q, bx->p, bx->g, bx->e, bx->ek, bx->mass, bx->tfwd, bx->tinv };
    for (double *b : bufs)
        if (b) (void)hipFree(b);
    if (bx->step_base) (void)hipFree(bx->step_base);
    if (bx->nhc) (void)hipFree(bx->nhc);
    if (bx->nhc_part) (void)hipFree(bx->nhc_part);
    if (bx->tr) (void)hipFree(bx->tr);
    if (bx->ev0) (void)hipEventDestroy(bx->ev0);
    if (bx->ev1) (void)hipEventDestroy(bx->ev1);
    if (bx->stream) (void)hipStreamDestroy(bx->stream);
    delete bx;
    return CAM_OK;
}

int cam_box_set_params(cam_box *bx, const double *theta)
{
    CAM_CHECK_ARG(bx && theta, "null argument");
    std::memcpy(bx->theta, theta, sizeof bx->theta);
    for (auto &kv : bx->graphs) (void)hipGraphExecDestroy(kv.second);   // parameters are baked into the graphs
    bx->graphs.clear();
    clear_error();
    if (bx->have_state) {
        Theta th;
        std::memcpy(th.p, bx->theta, sizeof th.p);
        int rc = bx->precision == CAM_PREC_MIXED
                     ? ff_launch<float>(bx->ff, th, bx->q, bx->nb, bx->e, bx->g, nullptr, nullptr, nullptr, bx->stream)
                     : ff_launch<double>(bx->ff, th, bx->q, bx->nb, bx->e, bx->g, nullptr, nullptr, nullptr, bx->stream);
        if (rc != CAM_OK) return rc;
        CAM_HIP(hipStreamSynchronize(bx->stream));
    }
    return CAM_OK;
}

int cam_box_set_state(cam_box *bx, const double *q, const double *p)
{
    CAM_CHECK_ARG(bx && q && p, "null argument");
    clear_error();
    const size_t n = (size_t)bx->natoms * 3 * bx->nb;
    CAM_HIP(hipMemcpyAsync(bx->q, q, n * sizeof(double), hipMemcpyHostToDevice, bx->stream));
    CAM_HIP(hipMemcpyAsync(bx->p, p, n * sizeof(double), hipMemcpyHostToDevice, bx->stream));
    bx->have_state = true;
    // a new state starts the Nose-Hoover chain at rest, as mdinit does
    if (bx->nhc) CAM_HIP(hipMemsetAsync(bx->nhc, 0, 5 * sizeof(double), bx->stream));
    return cam_box_set_params(bx, bx->theta);
}

int cam_box_set_transrot(cam_box *bx, int on)
{
    CAM_CHECK_ARG(bx, "null box");
    clear_error();
    if (on && !bx->tr) CAM_HIP(hipMalloc(&bx->tr, (16 + 15 * (size_t)kTrBlocks) * sizeof(double)));
    // the reference's total mass: every atom counted once per bead, then
    // times nbeads (src/transrot.f90:62-72)
    std::vector<double> m(bx->natoms);
    CAM_HIP(hipMemcpy(m.data(), bx->mass, m.size() * sizeof(double), hipMemcpyDeviceToHost));
    double tm = 0.0;
    for (int i = 0; i < bx->natoms; ++i)
        for (int k = 0; k < bx->nb; ++k) tm = tm + m[i];
    bx->totmass = tm * bx->nb;
    bx->transrot = on != 0;
    for (auto &kv : bx->graphs) (void)hipGraphExecDestroy(kv.second);   // part of the captured step
    bx->graphs.clear();
    return CAM_OK;
}

int cam_box_set_nhc(cam_box *bx, double nose_q)
{
    CAM_CHECK_ARG(bx, "null box");
    CAM_CHECK_ARG(nose_q >= 0.0 && std::isfinite(nose_q), "nose_q must be >= 0");
    CAM_CHECK_ARG(nose_q == 0.0 || bx->andersen_step == 0, "the Nose-Hoover chain replaces Andersen: andersen_step must be 0");
    clear_error();
    if (nose_q > 0.0 && !bx->nhc) {
        CAM_HIP(hipMalloc(&bx->nhc, 5 * sizeof(double)));
        CAM_HIP(hipMalloc(&bx->nhc_part, kBoxNhcBlocks * sizeof(double)));
    }
    if (bx->nhc) CAM_HIP(hipMemsetAsync(bx->nhc, 0, 5 * sizeof(double), bx->stream));
    CAM_HIP(hipStreamSynchronize(bx->stream));
    bx->nose_q = nose_q;
    for (auto &kv : bx->graphs) (void)hipGraphExecDestroy(kv.second);   // the chain is part of the captured step
    bx->graphs.clear();
    return CAM_OK;
}

int cam_box_get_state(cam_box *bx, double *q, double *p)
{
    CAM_CHECK_ARG(bx && (q || p), "null argument");
    CAM_CHECK_ARG(bx->have_state, "no state uploaded");
    clear_error();
    const size_t n = (size_t)bx->natoms * 3 * bx->nb;
    if (q) CAM_HIP(hipMemcpyAsync(q, bx->q, n * sizeof(double), hipMemcpyDeviceToHost, bx->stream));
    if (p) CAM_HIP(hipMemcpyAsync(p, bx->p, n * sizeof(double), hipMemcpyDeviceToHost, bx->stream));
    CAM_HIP(hipStreamSynchronize(bx->stream));
    return CAM_OK;
}

static int box_graph(cam_box *bx, int nsteps, hipGraphExec_t *out)
{
    auto it = bx->graphs.find(nsteps);
    if (it == bx->graphs.end()) {
        CAM_HIP(hipStreamBeginCapture(bx->stream, hipStreamCaptureModeThreadLocal));
        int rc = box_enqueue(bx, nsteps);
        hipGraph_t graph = nullptr;
        hipError_t ce = hipStreamEndCapture(bx->stream, &graph);
        if (rc != CAM_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        CAM_HIP(ce);
        hipGraphExec_t exec;
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        CAM_HIP(ie);
        it = bx->graphs.emplace(nsteps, exec).first;
    }
    *out = it->second;
    return CAM_OK;
}

int cam_box_prepare(cam_box *bx, int nsteps)
{
    CAM_CHECK_ARG(bx && nsteps >= 1, "bad argument");
    CAM_CHECK_ARG(bx->have_state, "no state uploaded");
    clear_error();
    hipGraphExec_t g;
    return box_graph(bx, nsteps, &g);
}

int cam_box_run(cam_box *bx, int nsteps)
{
    CAM_CHECK_ARG(bx && nsteps >= 0, "bad argument");
    CAM_CHECK_ARG(bx->have_state, "no state uploaded");
    clear_error();
    if (nsteps == 0) return CAM_OK;
    hipGraphExec_t g;
    int rc = box_graph(bx, nsteps, &g);
    if (rc != CAM_OK) return rc;
    CAM_HIP(hipEventRecord(bx->ev0, bx->stream));
    CAM_HIP(hipGraphLaunch(g, bx->stream));
    CAM_HIP(hipEventRecord(bx->ev1, bx->stream));
    bx->last_ms = -1.0;
    return CAM_OK;
}

int cam_box_sync(cam_box *bx, double *ms)
{
    CAM_CHECK_ARG(bx, "null argument");
    clear_error();
    CAM_HIP(hipStreamSynchronize(bx->stream));
    if (bx->last_ms < 0.0) {
        float f = 0.f;
        if (hipEventElapsedTime(&f, bx->ev0, bx->ev1) == hipSuccess) bx->last_ms = f;
    }
    if (ms) *ms = bx->last_ms;
    return CAM_OK;
}

int cam_box_energies(cam_box *bx, double *epot_beads, double *ekin, double *espring)
{
    CAM_CHECK_ARG(bx && epot_beads && ekin && espring, "null argument");
    CAM_CHECK_ARG(bx->have_state, "no state uploaded");
    clear_error();
    hipLaunchKernelGGL(box_energy_kernel, dim3(1), dim3(256), 0, bx->stream, bx->q, bx->p, bx->mass, bx->natoms, bx->nb,
                       bx->beta, bx->ek, bx->ek + 1);
    CAM_LAUNCH_CHECK();
    CAM_HIP(hipMemcpyAsync(epot_beads, bx->e, bx->nb * sizeof(double), hipMemcpyDeviceToHost, bx->stream));
    CAM_HIP(hipMemcpyAsync(ekin, bx->ek, sizeof(double), hipMemcpyDeviceToHost, bx->stream));
    CAM_HIP(hipMemcpyAsync(espring, bx->ek + 1, sizeof(double), hipMemcpyDeviceToHost, bx->stream));
    CAM_HIP(hipStreamSynchronize(bx->stream));
    return CAM_OK;
}

int cam_box_device_state(cam_box *bx, double **q, double **p, double **dvdq, void **stream)
{
    CAM_CHECK_ARG(bx, "null argument");
    if (q) *q = bx->q;
    if (p) *p = bx->p;
    if (dvdq) *dvdq = bx->g;
    if (stream) *stream = (void *)bx->stream;
    return CAM_OK;
}

}  // extern "C"
