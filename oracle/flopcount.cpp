// oracle/flopcount.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// Counts the algorithmic work of one bead-force-evaluation by compiling the
// oracle's surfaces with a counting number type in place of double: every
// +, -, *, / is one flop; elementary functions (exp, log, pow, sqrt, tanh,
// cosh, acos, sin, cos) are counted separately.  Prints JSON.
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
CD_FN(exp) CD_FN(log) CD_FN(sqrt) CD_FN(tanh) CD_FN(cosh) CD_FN(acos) CD_FN(sin) CD_FN(cos)
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
#undef double

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
    std::printf("{\"h3\": {\"flops\": %ld, \"transcendentals\": %ld}, "
                "\"ch4h\": {\"flops\": %ld, \"transcendentals\": %ld}}\n",
                h3f, h3t, g_flops, g_trans);
    return 0;
}
