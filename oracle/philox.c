/*
 * oracle/philox.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Philox-4x32-10 (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11) and Box-Muller
 * normals, restated from the published algorithm.  The reference draws its
 * Andersen momenta from the compiler's random_number (src/andersen.f90:121-126),
 * whose stream no test of the reference pins; the engine replaces it with
 * this counter-based stream, and the oracle restates it so both produce the
 * same momenta.  Pinned by the published known-answer vectors
 * (tests/test_oracle_rpmd.py).
 */
#include <math.h>
#include "oracle.h"

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Standard normal for (seed, traj, draw, coord = 3*atom + dim, bead). */
double oracle_andersen_normal(uint64_t seed, uint32_t traj, uint64_t draw, int coord, int bead)
{
    const uint32_t ctr[4] = { (uint32_t)bead | ((uint32_t)(coord >> 1) << 16), traj, (uint32_t)draw,
                              (uint32_t)(draw >> 32) ^ 0xA5A5u };
    const uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint32_t r[4];
    oracle_philox4x32_10(ctr, key, r);
    const double u1 = ((double)((((uint64_t)r[0] << 32) | r[1]) >> 11) + 1.0) / 9007199254740992.0;
    const double u2 = ((double)((((uint64_t)r[2] << 32) | r[3]) >> 11) + 1.0) / 9007199254740992.0;
    const double rad = sqrt(-2.0 * log(u1));
    const double ang = 6.283185307179586 * u2;
    return (coord & 1) ? rad * sin(ang) : rad * cos(ang);
}
