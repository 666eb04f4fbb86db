// Prints the coefficient tables of erfc_tables.hpp (a host program, not part
// of the library): the Chebyshev interpolant of erfcx(x) = exp(x^2) erfc(x)
// on [0, X] at 96 nodes in long double, c_k = (2 - [k = 0]) / N sum_j f(x_j)
// cos(k theta_j), truncated, then written in powers of t = 2 x / X - 1
// (T_{k+1} = 2 t T_k - T_{k-1}); on stderr the largest relative error of the
// FP64 tables evaluated by Horner in FP64.
//     g++ -O2 erfc_tables_gen.cpp -o gen && ./gen
#include <cmath>
#include <cstdio>
constexpr int kT = 40;
void fit(long double X, int nt, long double *mono)
{
    const int N = 96;
    const long double pi = 3.14159265358979323846264338327950288L;
    long double fv[N], th[N];
    for (int j = 0; j < N; ++j) {
        th[j] = pi * (j + 0.5L) / N;
        const long double xx = 0.5L * X * (std::cos(th[j]) + 1.0L);
        fv[j] = std::exp(xx * xx) * std::erfc(xx);
    }
    long double cheb[kT] = {}, tk[kT] = {}, tkm[kT] = {};
    for (int q = 0; q < kT; ++q) mono[q] = 0.0L;
    for (int k = 0; k < nt; ++k) {
        long double sum = 0.0L;
        for (int j = 0; j < N; ++j) sum += fv[j] * std::cos(k * th[j]);
        cheb[k] = (k == 0 ? 1.0L : 2.0L) * sum / N;
    }
    tkm[0] = 1.0L;
    tk[1] = 1.0L;
    mono[0] += cheb[0];
    for (int q = 0; q < kT; ++q) mono[q] += cheb[1] * tk[q];
    for (int k = 2; k < nt; ++k) {
        long double tn[kT] = {};
        for (int q = 0; q + 1 < kT; ++q) tn[q + 1] += 2.0L * tk[q];
        for (int q = 0; q < kT; ++q) tn[q] -= tkm[q];
        for (int q = 0; q < kT; ++q) {
            tkm[q] = tk[q];
            tk[q] = tn[q];
            mono[q] += cheb[k] * tn[q];
        }
    }
}
int main()
{
    struct { long double X; int nt; const char *name; bool f32; } cs[] = {
        {2.5L, 24, "kErfcNarrow64", false}, {2.5L, 16, "kErfcNarrow32", true},
        {6.0L, 36, "kErfcWide64", false}, {6.0L, 24, "kErfcWide32", true}};
    for (auto &c : cs) {
        long double m[kT];
        fit(c.X, c.nt, m);
        std::printf("constexpr %s %s[%d] = {\n", c.f32 ? "float" : "double", c.name, c.nt);
        for (int q = 0; q < c.nt; ++q) {
            if (c.f32) std::printf("    %af,\n", (double)(float)m[q]);
            else std::printf("    %a,\n", (double)m[q]);
        }
        std::printf("};\n");
    }
    // accuracy of the FP64 tables against erfcx in long double
    for (auto &c : cs) {
        if (c.f32) continue;
        long double m[kT];
        fit(c.X, c.nt, m);
        double mm[kT];
        for (int q = 0; q < c.nt; ++q) mm[q] = (double)m[q];
        double worst = 0;
        for (int i = 0; i <= 200000; ++i) {
            const double x = (double)c.X * i / 200000.0;
            const double t = std::fma(x, (double)(2.0L / c.X), -1.0);
            double h = mm[c.nt - 1];
            for (int k = c.nt - 2; k >= 0; --k) h = std::fma(h, t, mm[k]);
            const long double ref = std::exp((long double)x * x) * std::erfc((long double)x);
            worst = std::fmax(worst, (double)std::fabs((h - ref) / ref));
        }
        std::fprintf(stderr, "%s: max rel error of erfcx %.2e\n", c.name, worst);
    }
}
