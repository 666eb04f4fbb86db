// Counter-based normal deviates: Philox-4x32-10 (Salmon et al., SC'11) and
// the Box-Muller transform.  Used for the Andersen momenta so every
// (trajectory, draw, coordinate, bead) gets a reproducible deviate no matter
// how the work is scheduled.  Host and device share this header.
#pragma once
#include <cstdint>
#include <cmath>
#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define CAM_HD __host__ __device__ __forceinline__
#else
#define CAM_HD inline
#endif

namespace cam {

struct u32x4 { uint32_t x, y, z, w; };

CAM_HD void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo)
{
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

CAM_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo(M0, c.x, hi0, lo0);
        mulhilo(M1, c.z, hi1, lo1);
        c = u32x4{ hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0 };
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Uniform double in (0, 1] from 64 random bits (53 significant).
CAM_HD double u01_open0(uint32_t hi, uint32_t lo)
{
    const uint64_t m = (((uint64_t)hi << 32) | lo) >> 11;
    return ((double)m + 1.0) * (1.0 / 9007199254740992.0);
}

// Two standard normal deviates for counter (c0, c1, c2, c3) under key seed.
CAM_HD void normal2(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, double& z0, double& z1)
{
    const u32x4 r = philox4x32_10(u32x4{ c0, c1, c2, c3 }, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double u1 = u01_open0(r.x, r.y), u2 = u01_open0(r.z, r.w);
    const double rad = sqrt(-2.0 * log(u1));
    const double ang = 6.283185307179586 * u2;
    z0 = rad * cos(ang);
    z1 = rad * sin(ang);
}

// Deviate for Andersen draw 'draw' of trajectory 'traj', coordinate 'coord'
// (atom*3 + dim) and bead 'bead'.  Coordinates pair up (2m, 2m+1) on one
// Philox call.
CAM_HD double andersen_normal(uint64_t seed, uint32_t traj, uint64_t draw, int coord, int bead)
{
    double z0, z1;
    normal2(seed, (uint32_t)bead | ((uint32_t)(coord >> 1) << 16), traj, (uint32_t)draw,
            (uint32_t)(draw >> 32) ^ 0xA5A5u, z0, z1);
    return (coord & 1) ? z1 : z0;
}

}  // namespace cam
