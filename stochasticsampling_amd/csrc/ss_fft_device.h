// ss_fft_device.h -- radix stages of the LDS Stockham FFT, shared by the
// axis-pass kernels (ss_fft.hip) and the fused spectral-step kernels
// (ss_spectral.h), so that both run the same arithmetic on each line.
//
// Element e of line l lives at buf[l * ls + e * es].  Convention of FFTW
// (crates/fftw3/src/fft.rs:156-231): sign -1 forward, unnormalised.
#pragma once
#include "ss_device.h"

namespace ss {

// Products and sums are contracted into FMAs only within one expression
// (contract(on), fixed by the source), not at the compiler's choice across
// statements: the axis-pass and fused kernels instantiate these functions with
// run-time and compile-time sizes and must round alike.
template <typename T>
__device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
#pragma clang fp contract(on)
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Radix-R DFT of v in place; root(t) = exp(sign 2 pi i t / R) read from the
// twiddle table of the full length n at index t*(n/R).
template <typename T, int R>
__device__ inline void dft_small(cplx<T>* v, const cplx<T>* tw, int n, int sign) {
#pragma clang fp contract(on)
    if constexpr (R == 2) {
        cplx<T> a = v[0], b = v[1];
        v[0] = {a.x + b.x, a.y + b.y};
        v[1] = {a.x - b.x, a.y - b.y};
    } else if constexpr (R == 4) {
        cplx<T> a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
        cplx<T> s02 = {a0.x + a2.x, a0.y + a2.y}, d02 = {a0.x - a2.x, a0.y - a2.y};
        cplx<T> s13 = {a1.x + a3.x, a1.y + a3.y}, d13 = {a1.x - a3.x, a1.y - a3.y};
        // multiply d13 by (sign * i)
        cplx<T> jd13 = sign < 0 ? cplx<T>{d13.y, -d13.x} : cplx<T>{-d13.y, d13.x};
        v[0] = {s02.x + s13.x, s02.y + s13.y};
        v[2] = {s02.x - s13.x, s02.y - s13.y};
        v[1] = {d02.x + jd13.x, d02.y + jd13.y};
        v[3] = {d02.x - jd13.x, d02.y - jd13.y};
    } else {
        cplx<T> w[R];
        const int step = n / R;
#pragma unroll
        for (int q = 0; q < R; ++q) {
            T sr = T(0), si = T(0);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cplx<T> c = tw[((r * q) % R) * step];
                if (sign > 0) c.y = -c.y;
                sr += v[r].x * c.x - v[r].y * c.y;
                si += v[r].x * c.y + v[r].y * c.x;
            }
            w[q] = {sr, si};
        }
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = w[q];
    }
}

// One radix-R Stockham stage over nlines lines of length n (Ns = product of
// the radices of the earlier stages) by a workgroup of NT threads.
// LINES_FAST: consecutive threads take consecutive lines (lines interleaved in
// memory, es > 1); otherwise consecutive butterflies of one line (contiguous
// lines, es = 1).  NC > 0 fixes n = NC and Ns = NSC at compile time, NLC > 0
// the number of lines (the same arithmetic, with the index divisions folded).
// Butterfly b of a radix-R stage: line l, position j, twiddle index k.
template <bool LINES_FAST>
__device__ inline void butterfly_index(int b, int nlines, int m, int Ns, int& l, int& j, int& k) {
    if (LINES_FAST) {
        l = b % nlines;
        j = b / nlines;
    } else {
        l = b / m;
        j = b - l * m;
    }
    k = j % Ns;
}

// Reads butterfly (l, j, k) of line base s and transforms it in v.
template <typename T, int R>
__device__ inline void butterfly(const cplx<T>* s, const cplx<T>* tw, int n, int m, int tstep, int j, int k, int es,
                                 int sign, cplx<T> v[R]) {
#pragma clang fp contract(on)
#pragma unroll
    for (int r = 0; r < R; ++r) {
        cplx<T> x = s[(j + r * m) * es];
        if (r > 0 && k > 0) {
            cplx<T> c = tw[(r * k * tstep) % n];
            if (sign > 0) c.y = -c.y;
            x = cmul(x, c);
        }
        v[r] = x;
    }
    dft_small<T, R>(v, tw, n, sign);
}

template <typename T, int R, bool LINES_FAST, int NT, int NC = 0, int NSC = 0, int NLC = 0>
__device__ inline void stage(const cplx<T>* src, cplx<T>* dst, const cplx<T>* tw, int n_, int Ns_, int nlines_,
                             int ls, int es, int sign) {
    const int n = NC ? NC : n_;
    const int Ns = NC ? NSC : Ns_;
    const int nlines = NLC ? NLC : nlines_;
    const int m = n / R;
    const int tstep = n / (Ns * R);
    for (int b = threadIdx.x; b < nlines * m; b += NT) {
        int l, j, k;
        butterfly_index<LINES_FAST>(b, nlines, m, Ns, l, j, k);
        cplx<T> v[R];
        butterfly<T, R>(src + l * ls, tw, n, m, tstep, j, k, es, sign, v);
        const int idxD = (j / Ns) * Ns * R + k;
        cplx<T>* d = dst + l * ls;
#pragma unroll
        for (int r = 0; r < R; ++r) d[(idxD + r * Ns) * es] = v[r];
    }
}

// The same stage in place, for sizes fixed at compile time: every thread
// keeps its butterflies in registers between a read phase and a write phase
// (one LDS buffer instead of two).  Ends with a barrier.
template <typename T, int R, bool LINES_FAST, int NT, int N, int NS, int NL>
__device__ inline void stage_inplace(cplx<T>* buf, const cplx<T>* tw, int ls, int es, int sign) {
    constexpr int m = N / R, tstep = N / (NS * R), total = NL * m;
    constexpr int BPT = (total + NT - 1) / NT;
    cplx<T> v[BPT][R];
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        const int b = (int)threadIdx.x + u * NT;
        if (total % NT == 0 || b < total) {
            int l, j, k;
            butterfly_index<LINES_FAST>(b, NL, m, NS, l, j, k);
            butterfly<T, R>(buf + l * ls, tw, N, m, tstep, j, k, es, sign, v[u]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < BPT; ++u) {
        const int b = (int)threadIdx.x + u * NT;
        if (total % NT == 0 || b < total) {
            int l, j, k;
            butterfly_index<LINES_FAST>(b, NL, m, NS, l, j, k);
            const int idxD = (j / NS) * NS * R + k;
            cplx<T>* d = buf + l * ls;
#pragma unroll
            for (int r = 0; r < R; ++r) d[(idxD + r * NS) * es] = v[u][r];
        }
    }
    __syncthreads();
}

// Radix plan of one axis (factors 4, 2, 3, 5 only).
struct SpecAxis {
    int n, nf;
    int fac[16];
};

// All stages of one axis, ping-ponging between a and b; returns the buffer
// holding the result.  Ends with a barrier.
template <typename T, bool LINES_FAST, int NT>
__device__ inline cplx<T>* fft_lines(cplx<T>* a, cplx<T>* b, const cplx<T>* tw, const SpecAxis& ax, int nlines,
                                     int ls, int es, int sign) {
    cplx<T>* src = a;
    cplx<T>* dst = b;
    int Ns = 1;
    for (int s = 0; s < ax.nf; ++s) {
        const int R = ax.fac[s];
        switch (R) {
            case 4: stage<T, 4, LINES_FAST, NT>(src, dst, tw, ax.n, Ns, nlines, ls, es, sign); break;
            case 2: stage<T, 2, LINES_FAST, NT>(src, dst, tw, ax.n, Ns, nlines, ls, es, sign); break;
            case 3: stage<T, 3, LINES_FAST, NT>(src, dst, tw, ax.n, Ns, nlines, ls, es, sign); break;
            default: stage<T, 5, LINES_FAST, NT>(src, dst, tw, ax.n, Ns, nlines, ls, es, sign); break;
        }
        __syncthreads();
        cplx<T>* t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }
    return src;
}

// The radix of the next stage of factorize()'s plan when m points remain:
// the 4s first, then a 2, then the 3s and the 5s.
constexpr int next_radix(int m) { return m % 4 == 0 ? 4 : m % 2 == 0 ? 2 : m % 3 == 0 ? 3 : 5; }

// fft_lines for a length N known at compile time (factors 2, 3, 4 and 5, in
// the order of factorize()), over NL lines when NL > 0.
template <typename T, bool LINES_FAST, int NT, int N, int NL, int NS = 1>
__device__ inline cplx<T>* fft_lines_ct(cplx<T>* src, cplx<T>* dst, const cplx<T>* tw, int nlines, int ls, int es,
                                        int sign) {
    static_assert(N > 1, "length");
    if constexpr (NS >= N) {
        return src;
    } else {
        constexpr int R = next_radix(N / NS);
        static_assert((N / NS) % R == 0, "lengths with factors 2, 3 and 5 only");
        stage<T, R, LINES_FAST, NT, N, NS, NL>(src, dst, tw, N, NS, nlines, ls, es, sign);
        __syncthreads();
        return fft_lines_ct<T, LINES_FAST, NT, N, NL, NS * R>(dst, src, tw, nlines, ls, es, sign);
    }
}

// fft_lines_ct in place (NL lines, all sizes at compile time); the result is
// in buf.
template <typename T, bool LINES_FAST, int NT, int N, int NL, int NS = 1>
__device__ inline void fft_lines_ct_inplace(cplx<T>* buf, const cplx<T>* tw, int ls, int es, int sign) {
    if constexpr (NS < N) {
        constexpr int R = next_radix(N / NS);
        static_assert((N / NS) % R == 0, "lengths with factors 2, 3 and 5 only");
        stage_inplace<T, R, LINES_FAST, NT, N, NS, NL>(buf, tw, ls, es, sign);
        fft_lines_ct_inplace<T, LINES_FAST, NT, N, NL, NS * R>(buf, tw, ls, es, sign);
    }
}

// Whether the runtime plan of an axis is the one fft_lines_ct<N> runs.
inline bool ct_plan(const SpecAxis& a, int N) {
    if (a.n != N || N <= 1) return false;
    int m = N, k = 0;
    while (m > 1) {
        const int r = next_radix(m);
        if (m % r != 0 || k >= a.nf || a.fac[k] != r) return false;
        m /= r;
        ++k;
    }
    return k == a.nf;
}

}  // namespace ss
