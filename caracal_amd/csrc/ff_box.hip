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

// The j-range of split s of the nonbonded pair loop: every kernel that walks
// pairs by split uses the same ranges.
__host__ __device__ __forceinline__ int split_chunk(int n, int nsplit)
{
    return ((n + nsplit - 1) / nsplit + kTile - 1) / kTile * kTile;
}

// Per atom: the centroid of its beads (the mean of the minimum-image offsets
// from bead 0) and the largest distance of one of its beads from it.
__global__ void centroid_dev_kernel(int n, int nbeads, double box, const double *__restrict__ x,
                                    double *__restrict__ cen, double *__restrict__ dev)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x0[3], s[3] = { 0.0, 0.0, 0.0 };
    for (int c = 0; c < 3; ++c) x0[c] = x[3 * i + c];
    for (int b = 0; b < nbeads; ++b)
        for (int c = 0; c < 3; ++c) s[c] += min_image(x[((size_t)b * n + i) * 3 + c] - x0[c], box);
    double cc[3];
    for (int c = 0; c < 3; ++c) cc[c] = x0[c] + s[c] / nbeads;
    double m = 0.0;
    for (int b = 0; b < nbeads; ++b) {
        double r2 = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double d = min_image(x[((size_t)b * n + i) * 3 + c] - cc[c], box);
            r2 += d * d;
        }
        m = fmax(m, r2);
    }
    for (int c = 0; c < 3; ++c) cen[3 * i + c] = cc[c];
    dev[i] = sqrt(m);
}

// Neighbour list shared by all beads: atom i keeps every atom j of another
// molecule whose centroid lies within rc + dev_i + dev_j + margin of its own.
// For any bead, |d_ij| >= |c_i - c_j| - dev_i - dev_j, so no pair inside the
// cutoff in any bead is left out; margin covers the binary32 rounding of the
// mixed kernel's distances.  Row i lists j in increasing order; nofs[i][s] is
// where split s's j-range starts in it, so the pair loop meets the same pairs
// in the same order as the all-pairs loop.  One wave per atom.
__global__ __launch_bounds__(256) void nlist_kernel(FFDev f, int nsplit, double margin, const double *__restrict__ cen,
                                                    const double *__restrict__ dev, int *__restrict__ nbr,
                                                    int *__restrict__ nofs)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int n = f.natoms;
    if (i >= n) return;                    // the whole wave
    const double c0 = cen[3 * i], c1 = cen[3 * i + 1], c2 = cen[3 * i + 2], di = dev[i];
    const int mi = f.mol[i];
    const int chunk = split_chunk(n, nsplit);
    int *row = nbr + (size_t)i * n;
    int *ofs = nofs + (size_t)i * (nsplit + 1);
    int cnt = 0;
    for (int sp = 0; sp < nsplit; ++sp) {
        if (lane == 0) ofs[sp] = cnt;
        const int j0 = sp * chunk, j1 = min(n, j0 + chunk);
        for (int jb = j0; jb < j1; jb += 64) {
            const int j = jb + lane;
            bool keep = false;
            if (j < j1 && f.mol[j] != mi) {
                const double d0 = min_image(c0 - cen[3 * j], f.box), d1 = min_image(c1 - cen[3 * j + 1], f.box),
                             d2 = min_image(c2 - cen[3 * j + 2], f.box);
                const double lim = f.rc + di + dev[j] + margin;
                keep = d0 * d0 + d1 * d1 + d2 * d2 < lim * lim;
            }
            const unsigned long long m = __ballot(keep);
            if (keep) row[cnt + __popcll(m & ((1ull << lane) - 1ull))] = j;
            cnt += __popcll(m);
        }
    }
    if (lane == 0) ofs[nsplit] = cnt;
}

// Nonbonded pairs of i against a j-range; GRAD: gradient of i, PAR: parameter
// derivatives (and with v, the force-matching contraction).  LIST: the pairs
// come from the neighbour list (nbr, nofs) instead of all atoms of the range;
// the pairs left out are beyond the cutoff in every bead, so both forms add
// the same terms in the same order.
template <typename T, bool GRAD, bool PAR, bool LIST>
__global__ __launch_bounds__(kTile) void nonbonded_kernel(FFDev f, Theta th, const double *__restrict__ x,
                                                          const double *__restrict__ v, int nsplit,
                                                          double *__restrict__ gpart, double *__restrict__ epart,
                                                          const int *__restrict__ nbr, const int *__restrict__ nofs)
{
    __shared__ T sx[kTile][3];
    __shared__ T sv[kTile][3];
    __shared__ int st[kTile], sm[kTile];
    __shared__ double red[kTile];
    const int n = f.natoms;
    const int tid = threadIdx.x;
    const int i = blockIdx.x * kTile + tid;
    const int split = blockIdx.y, b = blockIdx.z;
    const double *xb = x + (size_t)b * n * 3;
    const double *vb = v ? v + (size_t)b * n * 3 : nullptr;
    const bool act = i < n;
    const int ii = act ? i : 0;
    const T xi0 = (T)xb[3 * ii], xi1 = (T)xb[3 * ii + 1], xi2 = (T)xb[3 * ii + 2];
    T vi0 = 0, vi1 = 0, vi2 = 0;
    if (PAR && vb) {
        vi0 = (T)vb[3 * ii];
        vi1 = (T)vb[3 * ii + 1];
        vi2 = (T)vb[3 * ii + 2];
    }
    const int ti = f.type[ii], mi = f.mol[ii];
    const T box = (T)f.box, rc2 = (T)(f.rc * f.rc), za = (T)f.za, zp = (T)f.zp, rc = (T)f.rc;
    const T two_over_spi = (T)1.1283791670955126;
    const T qi = (T)(ti == 0 ? th.p[5] : th.p[6]);
    // per-partner-type constants, selected by the partner's type below (no
    // run-time indexing of parameter arrays)
    const T qt[2] = { (T)th.p[5], (T)th.p[6] };
    const T c6t[2] = { (T)(ti == 0 ? th.p[7] : th.p[8]), (T)(ti == 0 ? th.p[8] : th.p[9]) };
    const T r0t[2] = { (T)(ti == 0 ? f.r0bj[0][0] : f.r0bj[1][0]), (T)(ti == 0 ? f.r0bj[0][1] : f.r0bj[1][1]) };
    const T srt[2] = { (T)(ti == 0 ? f.sr42[0][0] : f.sr42[1][0]), (T)(ti == 0 ? f.sr42[0][1] : f.sr42[1][1]) };
    const T zrt[2] = { (T)(ti == 0 ? f.zrep[0][0] : f.zrep[1][0]), (T)(ti == 0 ? f.zrep[0][1] : f.zrep[1][1]) };
    const T alt[2] = { (T)(ti == 0 ? f.arep[0][0] : f.arep[1][0]), (T)(ti == 0 ? f.arep[0][1] : f.arep[1][1]) };
    Acc<T> g0, g1, g2;
    double e = 0.0;
    // parameter derivatives: [0] q_O, [1] q_H, [2..4] C6 of (ti, partner)
    // pairs OO/OH/HH; dv* the same for the force-matching contraction
    double dq[2] = { 0.0, 0.0 }, dc[3] = { 0.0, 0.0, 0.0 }, dvq[2] = { 0.0, 0.0 }, dvc[3] = { 0.0, 0.0, 0.0 };

    // one pair (i, j): j's position, type and direction vector
    auto pair = [&](T xj0, T xj1, T xj2, int tj, T vj0, T vj1, T vj2) {
            const T d0 = min_image(xi0 - xj0, box);
            const T d1 = min_image(xi1 - xj1, box);
            const T d2 = min_image(xi2 - xj2, box);
            const T r2 = d0 * d0 + d1 * d1 + d2 * d2;
            if (r2 > rc2) return;
            const bool jo = tj == 0;
            const T r = sqrt_t(r2);
            const T rinv = (T)1 / r;
            const T qj = jo ? qt[0] : qt[1];
            const T c6 = jo ? c6t[0] : c6t[1];
            const T ev = erfc_t(za * r);
            const T phi = ev * rinv - zp * (r - rc);
            const T R0 = jo ? r0t[0] : r0t[1], sr = jo ? srt[0] : srt[1];
            const T R02 = R0 * R0, R06 = R02 * R02 * R02;
            const T r4 = r2 * r2, r6 = r4 * r2;
            const T t6 = r6 + R06, t8 = r6 * r2 + R06 * R02;
            const T it6 = (T)1 / t6, it8 = (T)1 / t8;
            const T dispu = -(it6 + sr * it8);
            const T zr = jo ? zrt[0] : zrt[1], al = jo ? alt[0] : alt[1];
            const T rep = zr * exp_t(-al * r) * rinv;
            e += 0.5 * (double)(qi * qj * phi + c6 * dispu + rep);
            if (GRAD || PAR) {
                const T dphi = -ev * rinv * rinv - two_over_spi * za * exp_t(-za * za * r2) * rinv - zp;
                const T ddisp = (T)6 * r4 * r * it6 * it6 + sr * (T)8 * r6 * r * it8 * it8;
                const T drep = -rep * (al + rinv);
                if (GRAD) {
                    const T fp = (qi * qj * dphi + c6 * ddisp + drep) * rinv;
                    g0.add(fp * d0);
                    g1.add(fp * d1);
                    g2.add(fp * d2);
                }
                if (PAR) {
                    // d/dq_type of q_i q_j: own charge gets q_j, partner's charge gets q_i
                    const double hj = 0.5 * (double)(qj * phi), hi = 0.5 * (double)(qi * phi);
                    if (ti == 0) dq[0] += hj; else dq[1] += hj;
                    if (jo) dq[0] += hi; else dq[1] += hi;
                    const int cp = ti + tj;
                    const double hd = 0.5 * (double)dispu;
                    if (cp == 0) dc[0] += hd; else if (cp == 1) dc[1] += hd; else dc[2] += hd;
                    if (vb) {
                        const T proj = ((vi0 - vj0) * d0 + (vi1 - vj1) * d1 + (vi2 - vj2) * d2) * rinv;
                        const double vj_ = 0.5 * (double)(qj * dphi * proj), vi_ = 0.5 * (double)(qi * dphi * proj);
                        if (ti == 0) dvq[0] += vj_; else dvq[1] += vj_;
                        if (jo) dvq[0] += vi_; else dvq[1] += vi_;
                        const double vd = 0.5 * (double)(ddisp * proj);
                        if (cp == 0) dvc[0] += vd; else if (cp == 1) dvc[1] += vd; else dvc[2] += vd;
                    }
                }
            }
    };
    const int chunk = split_chunk(n, nsplit);
    const int j0 = split * chunk, j1 = min(n, j0 + chunk);
    if constexpr (LIST) {
        if (act) {
            const int *row = nbr + (size_t)i * n;
            const int k0 = nofs[(size_t)i * (nsplit + 1) + split], k1 = nofs[(size_t)i * (nsplit + 1) + split + 1];
            for (int k = k0; k < k1; ++k) {
                const int j = row[k];
                T vj0 = 0, vj1 = 0, vj2 = 0;
                if (PAR && vb) {
                    vj0 = (T)vb[3 * j];
                    vj1 = (T)vb[3 * j + 1];
                    vj2 = (T)vb[3 * j + 2];
                }
                pair((T)xb[3 * j], (T)xb[3 * j + 1], (T)xb[3 * j + 2], f.type[j], vj0, vj1, vj2);
            }
        }
    } else {
        for (int jt = j0; jt < j1; jt += kTile) {
            __syncthreads();
            const int j = jt + tid;
            if (j < j1) {
                sx[tid][0] = (T)xb[3 * j];
                sx[tid][1] = (T)xb[3 * j + 1];
                sx[tid][2] = (T)xb[3 * j + 2];
                st[tid] = f.type[j];
                sm[tid] = f.mol[j];
                if (PAR && vb) {
                    sv[tid][0] = (T)vb[3 * j];
                    sv[tid][1] = (T)vb[3 * j + 1];
                    sv[tid][2] = (T)vb[3 * j + 2];
                }
            } else {
                sm[tid] = mi;            // padding: excluded as intramolecular
                st[tid] = 0;
                sx[tid][0] = sx[tid][1] = sx[tid][2] = 0;
                if (PAR) sv[tid][0] = sv[tid][1] = sv[tid][2] = 0;
            }
            __syncthreads();
            if (!act) continue;
            const int nt = min(kTile, j1 - jt);
            for (int jj = 0; jj < nt; ++jj) {
                if (sm[jj] == mi) continue;
                pair(sx[jj][0], sx[jj][1], sx[jj][2], st[jj], PAR ? sv[jj][0] : T(0), PAR ? sv[jj][1] : T(0),
                     PAR ? sv[jj][2] : T(0));
            }
        }
    }
    if (GRAD && act) {
        double *gp = gpart + (((size_t)split * gridDim.z + b) * n + i) * 3;
        gp[0] = g0.get();
        gp[1] = g1.get();
        gp[2] = g2.get();
    }
    // block sums of energy and parameter terms, fixed order
    const int nblk = gridDim.x * gridDim.y;
    const int blk = blockIdx.y * gridDim.x + blockIdx.x;
    double *ep = epart + ((size_t)b * (nblk + 1) + blk) * kNE;
    auto block_sum = [&](double val, int slot) {
        __syncthreads();
        red[tid] = act ? val : 0.0;
        __syncthreads();
        for (int w = kTile / 2; w > 0; w >>= 1) {
            if (tid < w) red[tid] += red[tid + w];
            __syncthreads();
        }
        if (tid == 0) ep[slot] = red[0];
    };
    block_sum(e, 0);
    if (PAR) {
        for (int p = 0; p < 5; ++p) {
            block_sum(0.0, 1 + p);
            block_sum(0.0, 1 + kNP + p);
        }
        block_sum(dq[0], 6);
        block_sum(dq[1], 7);
        block_sum(dc[0], 8);
        block_sum(dc[1], 9);
        block_sum(dc[2], 10);
        block_sum(dvq[0], 1 + kNP + 5);
        block_sum(dvq[1], 1 + kNP + 6);
        block_sum(dvc[0], 1 + kNP + 7);
        block_sum(dvc[1], 1 + kNP + 8);
        block_sum(dvc[2], 1 + kNP + 9);
    }
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
    __shared__ double red[kTile];
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
    const int blk = blockIdx.x;
    double *ep = epart + ((size_t)b * (nblk_nb + 1)) * kNE;   // bonded rows follow the nonbonded ones
    auto block_sum = [&](double val) -> double {
        __syncthreads();
        red[tid] = act ? val : 0.0;
        __syncthreads();
        for (int w = kTile / 2; w > 0; w >>= 1) {
            if (tid < w) red[tid] += red[tid + w];
            __syncthreads();
        }
        return red[0];
    };
    // bonded block sums go into a separate per-bead array slot region
    double *bp = epart + ((size_t)gridDim.y * (nblk_nb + 1)) * kNE + ((size_t)b * gridDim.x + blk) * kNE;
    (void)ep;
    double s0 = block_sum(e);
    if (tid == 0) bp[0] = s0;
    for (int p = 0; p < kNP; ++p) {
        const double sp = block_sum(dth[p]);
        if (tid == 0) bp[1 + p] = PAR ? sp : 0.0;
    }
    for (int p = 0; p < kNP; ++p) {
        const double sp = block_sum(dvth[p]);
        if (tid == 0) bp[1 + kNP + p] = PAR ? sp : 0.0;
    }
}

// Per-bead totals: nonbonded block rows then bonded block rows, fixed order.
__global__ void ff_reduce_kernel(const double *epart, int nblk_nb, int nblk_b, int nbeads, double *e, double *dth,
                                 double *dvth)
{
    const int b = blockIdx.x;
    const int slot = threadIdx.x;
    if (slot >= kNE) return;
    double s = 0.0;
    const double *nb = epart + (size_t)b * (nblk_nb + 1) * kNE;
    for (int k = 0; k < nblk_nb; ++k) s += nb[(size_t)k * kNE + slot];
    const double *bd = epart + ((size_t)nbeads * (nblk_nb + 1)) * kNE + (size_t)b * nblk_b * kNE;
    for (int k = 0; k < nblk_b; ++k) s += bd[(size_t)k * kNE + slot];
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
    int cap_beads = 0, nsplit = 4;
    int pair_mode = CAM_FF_PAIRS_LIST;
    int *nbr = nullptr, *nofs = nullptr;   // neighbour list [atom][atom], split offsets [atom][nsplit + 1]
    double *cen = nullptr, *dev = nullptr; // bead centroids and spreads of the last call
};

namespace {

// Largest box for which the neighbour list is kept: its rows hold up to
// natoms entries each (natoms^2 ints, 1 GiB at this size).
constexpr int kListMaxAtoms = 16384;

bool ff_uses_list(const cam_ff *ff)
{
    return ff->kind == 0 && ff->pair_mode == CAM_FF_PAIRS_LIST && ff->natoms <= kListMaxAtoms;
}

int ff_ensure(cam_ff *ff, int nbeads)
{
    if (ff->kind == 2) return CAM_OK;             // the QMDFF tables hold their own buffers
    if (ff_uses_list(ff) && !ff->nbr) {
        const size_t n = ff->natoms;
        CAM_HIP(hipMalloc(&ff->nbr, n * n * sizeof(int)));
        CAM_HIP(hipMalloc(&ff->nofs, n * (ff->nsplit + 1) * sizeof(int)));
        CAM_HIP(hipMalloc(&ff->cen, n * 3 * sizeof(double)));
        CAM_HIP(hipMalloc(&ff->dev, n * sizeof(double)));
    }
    if (nbeads <= ff->cap_beads) return CAM_OK;
    if (ff->gpart) CAM_HIP(hipFree(ff->gpart));
    if (ff->epart) CAM_HIP(hipFree(ff->epart));
    ff->gpart = ff->epart = nullptr;
    const int n = ff->natoms;
    const int nbx = (n + kTile - 1) / kTile;
    CAM_HIP(hipMalloc(&ff->gpart, (size_t)ff->nsplit * nbeads * n * 3 * sizeof(double)));
    const size_t rows = (size_t)nbeads * (nbx * ff->nsplit + 1) + (size_t)nbeads * nbx;
    CAM_HIP(hipMalloc(&ff->epart, rows * kNE * sizeof(double)));
    ff->cap_beads = nbeads;
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
    const int nblk_nb = nbx * ff->nsplit;
    const bool grad = g != nullptr;
    const bool par = dth != nullptr || dvth != nullptr;
    dim3 gnb(nbx, ff->nsplit, nbeads);
    if (ff->kind == 2) return qmdff_launch<T>(ff->qm, x, nbeads, e, g, s);
    if (ff->kind == 1) {
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
        hipLaunchKernelGGL(ff_reduce_kernel, dim3(nbeads), dim3(32), 0, s, ff->epart, nblk_nb, nbx, nbeads, e, nullptr,
                           nullptr);
        CAM_LAUNCH_CHECK();
        return CAM_OK;
    }
    if (ff_uses_list(ff)) {
        // the list of this call's beads: centroids and spreads, then one wave per atom
        hipLaunchKernelGGL(centroid_dev_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, nbeads, ff->d.box, x,
                           ff->cen, ff->dev);
        CAM_LAUNCH_CHECK();
        hipLaunchKernelGGL(nlist_kernel, dim3((n + 3) / 4), dim3(256), 0, s, ff->d, ff->nsplit, 1e-3, ff->cen,
                           ff->dev, ff->nbr, ff->nofs);
        CAM_LAUNCH_CHECK();
        if (grad && par)
            hipLaunchKernelGGL((nonbonded_kernel<T, true, true, true>), gnb, dim3(kTile), 0, s, ff->d, th, x, v,
                               ff->nsplit, ff->gpart, ff->epart, ff->nbr, ff->nofs);
        else if (grad)
            hipLaunchKernelGGL((nonbonded_kernel<T, true, false, true>), gnb, dim3(kTile), 0, s, ff->d, th, x,
                               nullptr, ff->nsplit, ff->gpart, ff->epart, ff->nbr, ff->nofs);
        else if (par)
            hipLaunchKernelGGL((nonbonded_kernel<T, false, true, true>), gnb, dim3(kTile), 0, s, ff->d, th, x, v,
                               ff->nsplit, ff->gpart, ff->epart, ff->nbr, ff->nofs);
        else
            hipLaunchKernelGGL((nonbonded_kernel<T, false, false, true>), gnb, dim3(kTile), 0, s, ff->d, th, x,
                               nullptr, ff->nsplit, ff->gpart, ff->epart, ff->nbr, ff->nofs);
    } else {
        const int *nl = nullptr;
        if (grad && par)
            hipLaunchKernelGGL((nonbonded_kernel<T, true, true, false>), gnb, dim3(kTile), 0, s, ff->d, th, x, v,
                               ff->nsplit, ff->gpart, ff->epart, nl, nl);
        else if (grad)
            hipLaunchKernelGGL((nonbonded_kernel<T, true, false, false>), gnb, dim3(kTile), 0, s, ff->d, th, x,
                               nullptr, ff->nsplit, ff->gpart, ff->epart, nl, nl);
        else if (par)
            hipLaunchKernelGGL((nonbonded_kernel<T, false, true, false>), gnb, dim3(kTile), 0, s, ff->d, th, x, v,
                               ff->nsplit, ff->gpart, ff->epart, nl, nl);
        else
            hipLaunchKernelGGL((nonbonded_kernel<T, false, false, false>), gnb, dim3(kTile), 0, s, ff->d, th, x,
                               nullptr, ff->nsplit, ff->gpart, ff->epart, nl, nl);
    }
    CAM_LAUNCH_CHECK();
    dim3 gb(nbx, nbeads);
    if (grad && par)
        hipLaunchKernelGGL((bonded_kernel<true, true>), gb, dim3(kTile), 0, s, ff->d, th, x, v, ff->nsplit, ff->gpart, g,
                           ff->epart, nblk_nb);
    else if (grad)
        hipLaunchKernelGGL((bonded_kernel<true, false>), gb, dim3(kTile), 0, s, ff->d, th, x, nullptr, ff->nsplit,
                           ff->gpart, g, ff->epart, nblk_nb);
    else if (par)
        hipLaunchKernelGGL((bonded_kernel<false, true>), gb, dim3(kTile), 0, s, ff->d, th, x, v, ff->nsplit, ff->gpart,
                           g, ff->epart, nblk_nb);
    else
        hipLaunchKernelGGL((bonded_kernel<false, false>), gb, dim3(kTile), 0, s, ff->d, th, x, nullptr, ff->nsplit,
                           ff->gpart, g, ff->epart, nblk_nb);
    CAM_LAUNCH_CHECK();
    hipLaunchKernelGGL(ff_reduce_kernel, dim3(nbeads), dim3(32), 0, s, ff->epart, nblk_nb, nbx, nbeads, e, dth, dvth);
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
    int *lb[] = { ff->nbr, ff->nofs };
    for (int *p : lb)
        if (p) (void)hipFree(p);
    if (ff->cen) (void)hipFree(ff->cen);
    if (ff->dev) (void)hipFree(ff->dev);
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

__global__ void box_nhc_kernel(const double *part, int nparts, double *st, double beta, double nose_q, int natoms,
                               int nb, double dt)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double ek = 0.0;
    for (int i = 0; i < nparts; ++i) ek += part[i];
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
