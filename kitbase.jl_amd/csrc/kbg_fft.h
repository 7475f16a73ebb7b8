// kbg_fft.h -- forward DFT of a short sequence held in registers, for the pencils of the fast
// spectral collision's 3D transforms (FFTW's forward sign: X_k = sum_j x_j exp(-2 pi i j k / N)).
//
// Mixed-radix decimation in time, unrolled at compile time: N = R M with R = 4, 3 or 2, the R
// decimated subsequences transformed recursively, then X[k + s M] = sum_q W_N^{q k} W_R^{q s} Y_q[k].
// Every index is a compile-time constant once the recursion is inlined, so the sequences live in
// registers; twiddles W_N^j that are 1, -1, -i or +i are applied exactly, the others come from a
// table tw[j] = exp(-2 pi i j / NT) of the top-level length NT (uniform loads).
#pragma once

#include <hip/hip_runtime.h>

namespace kbg {

#define KBG_HD __host__ __device__ __forceinline__

constexpr int fft_radix(int n) { return n % 4 == 0 ? 4 : (n % 3 == 0 ? 3 : 2); }
constexpr bool fft_supported(int n) {
    while (n % 4 == 0) n /= 4;
    while (n % 3 == 0) n /= 3;
    while (n % 2 == 0) n /= 2;
    return n == 1;
}

template <int NT> KBG_HD float2 fft_twiddle(float2 a, int j, const float2* __restrict__ tw) {
    if (j == 0) return a;
    if (4 * j == NT) return make_float2(a.y, -a.x);       // W = -i
    if (2 * j == NT) return make_float2(-a.x, -a.y);      // W = -1
    if (4 * j == 3 * NT) return make_float2(-a.y, a.x);   // W = +i
    const float2 w = tw[j];
    return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

// Z[s] <- sum_q Z[q] exp(-2 pi i q s / R)
template <int R> KBG_HD void fft_butterfly(float2* Z);
template <> KBG_HD void fft_butterfly<2>(float2* Z) {
    const float2 a = Z[0], b = Z[1];
    Z[0] = make_float2(a.x + b.x, a.y + b.y);
    Z[1] = make_float2(a.x - b.x, a.y - b.y);
}
template <> KBG_HD void fft_butterfly<3>(float2* Z) {
    // W = exp(-2 pi i / 3) = -1/2 - i sqrt(3)/2:  X1,2 = Z0 - (Z1 + Z2)/2 -+ i sqrt(3)/2 (Z1 - Z2)
    const float h = 0.866025403784438647f;
    const float2 s = make_float2(Z[1].x + Z[2].x, Z[1].y + Z[2].y);
    const float2 d = make_float2(Z[1].x - Z[2].x, Z[1].y - Z[2].y);
    const float2 m = make_float2(Z[0].x - 0.5f * s.x, Z[0].y - 0.5f * s.y);
    Z[0] = make_float2(Z[0].x + s.x, Z[0].y + s.y);
    Z[1] = make_float2(m.x + h * d.y, m.y - h * d.x);
    Z[2] = make_float2(m.x - h * d.y, m.y + h * d.x);
}
template <> KBG_HD void fft_butterfly<4>(float2* Z) {
    const float2 a = make_float2(Z[0].x + Z[2].x, Z[0].y + Z[2].y), b = make_float2(Z[0].x - Z[2].x, Z[0].y - Z[2].y);
    const float2 c = make_float2(Z[1].x + Z[3].x, Z[1].y + Z[3].y), d = make_float2(Z[1].x - Z[3].x, Z[1].y - Z[3].y);
    Z[0] = make_float2(a.x + c.x, a.y + c.y);
    Z[1] = make_float2(b.x + d.y, b.y - d.x);   // b - i d
    Z[2] = make_float2(a.x - c.x, a.y - c.y);
    Z[3] = make_float2(b.x - d.y, b.y + d.x);   // b + i d
}

// X[k] = sum_{j < N} x[j S] exp(-2 pi i j k / N), k < N; NT is the top-level length (NT / N an integer)
template <int N, int NT, int S> KBG_HD void fft_dit(const float2* x, float2* X, const float2* __restrict__ tw) {
    if constexpr (N == 1) {
        X[0] = x[0];
    } else {
        constexpr int R = fft_radix(N), M = N / R, TS = NT / N;
        float2 Y[N];
#pragma unroll
        for (int q = 0; q < R; q++) fft_dit<M, NT, S * R>(x + q * S, Y + q * M, tw);
#pragma unroll
        for (int k = 0; k < M; k++) {
            float2 Z[R];
#pragma unroll
            for (int q = 0; q < R; q++) Z[q] = fft_twiddle<NT>(Y[q * M + k], q * k * TS, tw);
            fft_butterfly<R>(Z);
#pragma unroll
            for (int s = 0; s < R; s++) X[k + s * M] = Z[s];
        }
    }
}

}  // namespace kbg
