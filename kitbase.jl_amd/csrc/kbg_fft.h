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

// complex values as float2 or double2
template <class C> struct FftReal;
template <> struct FftReal<float2> { using type = float; };
template <> struct FftReal<double2> { using type = double; };
template <class C> KBG_HD C fft_cx(typename FftReal<C>::type x, typename FftReal<C>::type y) {
    C z;
    z.x = x;
    z.y = y;
    return z;
}

template <int NT, class C> KBG_HD C fft_twiddle(C a, int j, const C* __restrict__ tw) {
    if (j == 0) return a;
    if (4 * j == NT) return fft_cx<C>(a.y, -a.x);       // W = -i
    if (2 * j == NT) return fft_cx<C>(-a.x, -a.y);      // W = -1
    if (4 * j == 3 * NT) return fft_cx<C>(-a.y, a.x);   // W = +i
    const C w = tw[j];
    return fft_cx<C>(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

// Z[s] <- sum_q Z[q] exp(-2 pi i q s / R)
template <int R, class C> struct FftButterfly;
template <class C> struct FftButterfly<2, C> {
    static KBG_HD void run(C* Z) {
        const C a = Z[0], b = Z[1];
        Z[0] = fft_cx<C>(a.x + b.x, a.y + b.y);
        Z[1] = fft_cx<C>(a.x - b.x, a.y - b.y);
    }
};
template <class C> struct FftButterfly<3, C> {
    // W = exp(-2 pi i / 3) = -1/2 - i sqrt(3)/2:  X1,2 = Z0 - (Z1 + Z2)/2 -+ i sqrt(3)/2 (Z1 - Z2)
    static KBG_HD void run(C* Z) {
        using R = typename FftReal<C>::type;
        const R h = R(0.86602540378443864676);
        const C s = fft_cx<C>(Z[1].x + Z[2].x, Z[1].y + Z[2].y);
        const C d = fft_cx<C>(Z[1].x - Z[2].x, Z[1].y - Z[2].y);
        const C m = fft_cx<C>(Z[0].x - R(0.5) * s.x, Z[0].y - R(0.5) * s.y);
        Z[0] = fft_cx<C>(Z[0].x + s.x, Z[0].y + s.y);
        Z[1] = fft_cx<C>(m.x + h * d.y, m.y - h * d.x);
        Z[2] = fft_cx<C>(m.x - h * d.y, m.y + h * d.x);
    }
};
template <class C> struct FftButterfly<4, C> {
    static KBG_HD void run(C* Z) {
        const C a = fft_cx<C>(Z[0].x + Z[2].x, Z[0].y + Z[2].y), b = fft_cx<C>(Z[0].x - Z[2].x, Z[0].y - Z[2].y);
        const C c = fft_cx<C>(Z[1].x + Z[3].x, Z[1].y + Z[3].y), d = fft_cx<C>(Z[1].x - Z[3].x, Z[1].y - Z[3].y);
        Z[0] = fft_cx<C>(a.x + c.x, a.y + c.y);
        Z[1] = fft_cx<C>(b.x + d.y, b.y - d.x);  // b - i d
        Z[2] = fft_cx<C>(a.x - c.x, a.y - c.y);
        Z[3] = fft_cx<C>(b.x - d.y, b.y + d.x);  // b + i d
    }
};

// X[k] = sum_{j < N} x[j S] exp(-2 pi i j k / N), k < N; NT is the top-level length (NT / N an integer)
template <int N, int NT, int S, class C> KBG_HD void fft_dit(const C* x, C* X, const C* __restrict__ tw) {
    if constexpr (N == 1) {
        X[0] = x[0];
    } else {
        constexpr int R = fft_radix(N), M = N / R, TS = NT / N;
        C Y[N];
#pragma unroll
        for (int q = 0; q < R; q++) fft_dit<M, NT, S * R, C>(x + q * S, Y + q * M, tw);
#pragma unroll
        for (int k = 0; k < M; k++) {
            C Z[R];
#pragma unroll
            for (int q = 0; q < R; q++) Z[q] = fft_twiddle<NT, C>(Y[q * M + k], q * k * TS, tw);
            FftButterfly<R, C>::run(Z);
#pragma unroll
            for (int s = 0; s < R; s++) X[k + s * M] = Z[s];
        }
    }
}

}  // namespace kbg
