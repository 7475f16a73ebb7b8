// Host build of kbg_fft.h's register FFT for tests/test_fft.py: reads "N precision" and N complex
// inputs (re im pairs) on stdin, writes the N outputs of the forward transform (FFTW's sign).
#include "../../kitbase.jl_amd/csrc/kbg_fft.h"

#include <cmath>
#include <cstdio>
#include <cstring>

using namespace kbg;

template <int N, class C> static void run() {
    C tw[N], x[N], X[N];
    for (int j = 0; j < N; j++) tw[j] = fft_cx<C>(std::cos(-2.0 * M_PI * j / N), std::sin(-2.0 * M_PI * j / N));
    for (int j = 0; j < N; j++) {
        double re, im;
        if (std::scanf("%lf %lf", &re, &im) != 2) return;
        x[j] = fft_cx<C>(re, im);
    }
    fft_dit<N, N, 1, C>(x, X, tw);
    for (int k = 0; k < N; k++) std::printf("%.17g %.17g\n", (double)X[k].x, (double)X[k].y);
}

template <class C> static int dispatch(int n) {
    switch (n) {
    case 2: run<2, C>(); return 0;
    case 3: run<3, C>(); return 0;
    case 4: run<4, C>(); return 0;
    case 6: run<6, C>(); return 0;
    case 8: run<8, C>(); return 0;
    case 12: run<12, C>(); return 0;
    case 16: run<16, C>(); return 0;
    case 18: run<18, C>(); return 0;
    case 24: run<24, C>(); return 0;
    default: return 1;
    }
}

int main() {
    int n;
    char prec[8];
    if (std::scanf("%d %7s", &n, prec) != 2) return 2;
    return std::strcmp(prec, "f32") == 0 ? dispatch<float2>(n) : dispatch<double2>(n);
}
