// oracle/flopcount.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Counts the algorithmic work of one bead-force-evaluation by compiling the
// oracle's surfaces (and the box force field, oracle/ff_box.c) with a
// counting number type in place of double: every +, -, *, / is one flop;
// elementary functions (exp, log, pow, sqrt, tanh, cosh, acos, sin, cos,
// erfc) are counted separately; rint (the minimum image) is one flop.  For the
// box it prints the work of one intermolecular pair inside and outside the
// cutoff (energy and gradient), of one bond and of one angle.  Prints JSON.
//   g++ -O1 -std=c++17 flopcount.cpp -o flopcount && ./flopcount
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <math.h>

static long g_flops = 0, g_trans = 0;

struct Cd {
    double v;
    Cd() : v(0) {}
    Cd(double x) : v(x) {}
    explicit operator float() const { return (float)v; }
    Cd& operator+=(Cd o) { ++g_flops; v += o.v; return *this; }
    Cd& operator-=(Cd o) { ++g_flops; v -= o.v; return *this; }
    Cd& operator*=(Cd o) { ++g_flops; v *= o.v; return *this; }
    Cd& operator/=(Cd o) { ++g_flops; v /= o.v; return *this; }
};
inline Cd operator+(Cd a, Cd b) { ++g_flops; return Cd(a.v + b.v); }
inline Cd operator-(Cd a, Cd b) { ++g_flops; return Cd(a.v - b.v); }
inline Cd operator*(Cd a, Cd b) { ++g_flops; return Cd(a.v * b.v); }
inline Cd operator/(Cd a, Cd b) { ++g_flops; return Cd(a.v / b.v); }
inline Cd operator-(Cd a) { return Cd(-a.v); }
inline bool operator<(Cd a, Cd b) { return a.v < b.v; }
inline bool operator>(Cd a, Cd b) { return a.v > b.v; }
inline bool operator<=(Cd a, Cd b) { return a.v <= b.v; }
inline bool operator>=(Cd a, Cd b) { return a.v >= b.v; }
inline bool operator==(Cd a, Cd b) { return a.v == b.v; }
#define CD_FN(name)                                  \
    inline Cd name(Cd a) { ++g_trans; return Cd(std::name(a.v)); }
CD_FN(exp) CD_FN(log) CD_FN(sqrt) CD_FN(tanh) CD_FN(cosh) CD_FN(acos) CD_FN(sin) CD_FN(cos) CD_FN(erfc)
inline Cd rint(Cd a) { ++g_flops; return Cd(std::rint(a.v)); }
inline Cd pow(Cd a, Cd b) { ++g_trans; return Cd(std::pow(a.v, b.v)); }
inline Cd fabs(Cd a) { return Cd(std::fabs(a.v)); }

// compile the oracle surfaces with Cd as their floating type
#define double Cd
#define SP(x) (Cd((double_sp_helper)(x)))
typedef float double_sp_helper;
#define oracle_h3_pote cnt_h3_pote
#define oracle_egrad_h3 cnt_egrad_h3
#define oracle_ch4h_pot cnt_ch4h_pot
#define oracle_egrad_ch4h cnt_egrad_ch4h
#define CARACAL_ORACLE_H
namespace h3ns {
#include "pes_h3.c"
}
namespace ch4hns {
#include "pes_ch4h.c"
}
namespace boxns {
#define ORACLE_FF_NPAR 10
typedef struct {
    int natoms, nbonds, nangles;
    const int *type, *mol, *bonds, *angles;
    double box, rcut, zahn_a, zahn_par;
    double sr42[2][2], r0bj[2][2], zrep[2][2], arep[2][2];
} oracle_ff;
#include "ff_box.c"
}
#undef double

// Work of one box force evaluation of natoms atoms at positions x.
static long box_flops(int natoms, const int *type, const int *mol, int nbonds, const int *bonds, int nangles,
                      const int *angles, const double *xd, long *trans)
{
    boxns::oracle_ff f{};
    f.natoms = natoms;
    f.nbonds = nbonds;
    f.nangles = nangles;
    f.type = type;
    f.mol = mol;
    f.bonds = bonds;
    f.angles = angles;
    f.box = 200.0;
    f.rcut = 18.897261;           // 10 angstrom
    f.zahn_a = 0.2 * 0.52917721092;
    f.zahn_par = 1e-4;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            f.sr42[a][b] = 2.7;
            f.r0bj[a][b] = 4.0;
            f.zrep[a][b] = 1.0;
            f.arep[a][b] = 1.5;
        }
    Cd th[ORACLE_FF_NPAR] = { 0.2, 1.2, 1.81, 0.15, 1.82, -0.8, 0.4, 10.0, 3.0, 1.0 };
    Cd x[27], g[27];
    for (int i = 0; i < 3 * natoms; ++i) x[i] = xd[i];
    g_flops = g_trans = 0;
    boxns::oracle_ff_eval(&f, th, x, g, nullptr, nullptr, nullptr);
    *trans = g_trans;
    return g_flops;
}

int main()
{
    // H3 at the collinear saddle, CH4H at its saddle region (representative branches)
    Cd qh[9] = { 0, 0, -1.757, 0.01, 0, 0, 0, 0.02, 1.757 };
    Cd v[1], g[18];
    g_flops = g_trans = 0;
    h3ns::cnt_egrad_h3(qh, 3, 1, v, g);
    const long h3f = g_flops, h3t = g_trans;
    Cd qc[18] = { -1.14, -1.14, 1.27, 0.008, 0.008, 0.008, 1.53, 1.52, 1.52,
                  1.27, -1.14, -1.14, -1.14, 1.27, -1.14, 2.48, 2.47, 2.48 };
    g_flops = g_trans = 0;
    ch4hns::cnt_egrad_ch4h(qc, 6, 1, v, g);
    const long chf = g_flops, cht = g_trans;
    // box: two atoms of different molecules 5 bohr apart (inside the cutoff)
    // and 30 bohr apart (outside), one bond, one angle; the empty system is
    // the fixed overhead to subtract
    const int ty[3] = { 0, 1, 1 }, m2[3] = { 0, 1, 2 }, m1[3] = { 0, 0, 0 };
    const int bd[2] = { 0, 1 }, an[3] = { 1, 0, 2 };
    const double xin[6] = { 0, 0, 0, 5.0, 0, 0 }, xout[6] = { 0, 0, 0, 30.0, 0, 0 };
    const double xang[9] = { 0, 0, 0, 1.8, 0, 0, -0.45, 1.74, 0 };
    long t0, tin, tout, tb, ta;
    const long f0 = box_flops(0, ty, m2, 0, bd, 0, an, xin, &t0);
    const long fin = box_flops(2, ty, m2, 0, bd, 0, an, xin, &tin) - f0;
    const long fout = box_flops(2, ty, m2, 0, bd, 0, an, xout, &tout) - f0;
    const long fb = box_flops(2, ty, m1, 1, bd, 0, an, xin, &tb) - f0;
    const long fa = box_flops(3, ty, m1, 0, bd, 1, an, xang, &ta) - f0;
    std::printf("{\"h3\": {\"flops\": %ld, \"transcendentals\": %ld}, "
                "\"ch4h\": {\"flops\": %ld, \"transcendentals\": %ld}, "
                "\"box\": {\"pair_in\": %ld, \"pair_in_transcendentals\": %ld, \"pair_out\": %ld, "
                "\"bond\": %ld, \"bond_transcendentals\": %ld, \"angle\": %ld, \"angle_transcendentals\": %ld}}\n",
                h3f, h3t, chf, cht, fin, tin, fout, fb, tb, fa, ta);
    return 0;
}
