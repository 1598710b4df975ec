// Batched radix FFT building blocks for CDNA4.
//
// A workgroup owns an LDS tile of B independent length-N complex sequences,
// element (n, b) at lds[n * LS + b] (LS >= B is the LDS row stride).  The
// transform is a Stockham autosort FFT: stage s with radix R after
// sub-length L computes, for every butterfly j in [0, N/R),
//     a_r = x[j + r N/R] * w_{L R}^{r k},   k = j mod L
//     x'[(j / L) L R + k + r L] = DFT_R(a)_r
// which leaves the result in natural order after the last stage.  Each
// stage reads its butterfly inputs from LDS into registers, synchronises,
// and writes the outputs back, so one tile buffer suffices.  Lanes are laid
// out with the batch index b fastest, so LDS accesses are unit-stride.
//
// Inter-stage twiddles w_N^t = exp(-2 pi i t / N) come from a table computed
// in double precision on the host and rounded to T.  Twiddles inside the
// small DFTs (R <= 32) are compile-time constants.
#pragma once
#include <type_traits>
#include "common.hpp"

namespace bfpsh {

// ---------------------------------------------------------------- helpers
template <int I, int N, class F>
BFPSH_DEV void sfor(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        sfor<I + 1, N>(f);
    }
}

// cos(2 pi q / 32), q in [0, 8]; other angles by symmetry.
constexpr double kCos32[9] = {
    1.0,
    0.98078528040323044912618223613424,
    0.92387953251128675612818318939679,
    0.83146961230254523707878837761791,
    0.70710678118654752440084436210485,
    0.55557023301960222474283081394853,
    0.38268343236508977172845998403040,
    0.19509032201612826784828486847702,
    0.0};
constexpr double cos32(int q) {
    q = ((q % 32) + 32) % 32;
    return q <= 8 ? kCos32[q] : q <= 16 ? -kCos32[16 - q] : q <= 24 ? -kCos32[q - 16] : kCos32[32 - q];
}
constexpr double sin32(int q) { return cos32(8 - q); }

// multiply by exp(SIGN 2 pi i M / R), R | 32, M a compile-time constant
template <typename T, int R, int M, int SIGN>
BFPSH_DEV C2<T> twc(C2<T> a) {
    constexpr int q = ((M % R) * (32 / R)) % 32;
    if constexpr (q == 0) {
        return a;
    } else if constexpr (q == 8) {
        return cmul_i<T, SIGN>(a);
    } else if constexpr (q == 16) {
        return mk<T>(-a.x, -a.y);
    } else if constexpr (q == 24) {
        return cmul_i<T, -SIGN>(a);
    } else {
        constexpr T c = T(cos32(q));
        constexpr T s = T(SIGN * sin32(q));
        return mk<T>(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

// ---------------------------------------------------------------- small DFTs
template <typename T, int R, int SIGN> struct Dft;

template <typename T, int SIGN> struct Dft<T, 2, SIGN> {
    static BFPSH_DEV void run(C2<T>* a) {
        C2<T> t = a[0];
        a[0] = cadd<T>(t, a[1]);
        a[1] = csub<T>(t, a[1]);
    }
};
template <typename T, int SIGN> struct Dft<T, 3, SIGN> {
    static BFPSH_DEV void run(C2<T>* a) {
        constexpr T h = T(0.86602540378443864676372317075294);  // sqrt(3)/2
        C2<T> t = cadd<T>(a[1], a[2]);
        C2<T> d = csub<T>(a[1], a[2]);
        C2<T> m = mk<T>(a[0].x - T(0.5) * t.x, a[0].y - T(0.5) * t.y);
        C2<T> r = cmul_i<T, SIGN>(mk<T>(h * d.x, h * d.y));
        a[0] = cadd<T>(a[0], t);
        a[1] = cadd<T>(m, r);
        a[2] = csub<T>(m, r);
    }
};
template <typename T, int SIGN> struct Dft<T, 4, SIGN> {
    static BFPSH_DEV void run(C2<T>* a) {
        C2<T> t0 = cadd<T>(a[0], a[2]), t1 = csub<T>(a[0], a[2]);
        C2<T> t2 = cadd<T>(a[1], a[3]), t3 = cmul_i<T, SIGN>(csub<T>(a[1], a[3]));
        a[0] = cadd<T>(t0, t2);
        a[2] = csub<T>(t0, t2);
        a[1] = cadd<T>(t1, t3);
        a[3] = csub<T>(t1, t3);
    }
};
// R = R1 * R2: X[k2 + R2 k1] = sum_n1 w_R1^{n1 k1} w_R^{n1 k2} sum_n2 a[n1 + R1 n2] w_R2^{n2 k2}
template <typename T, int R1, int R2, int SIGN>
BFPSH_DEV void dft_compose(C2<T>* a) {
    constexpr int R = R1 * R2;
    C2<T> c[R1][R2];
    sfor<0, R1>([&](auto n1) {
        C2<T> b[R2];
        sfor<0, R2>([&](auto n2) { b[n2] = a[n1 + R1 * n2]; });
        Dft<T, R2, SIGN>::run(b);
        sfor<0, R2>([&](auto k2) { c[n1][k2] = twc<T, R, n1 * k2, SIGN>(b[k2]); });
    });
    sfor<0, R2>([&](auto k2) {
        C2<T> d[R1];
        sfor<0, R1>([&](auto n1) { d[n1] = c[n1][k2]; });
        Dft<T, R1, SIGN>::run(d);
        sfor<0, R1>([&](auto k1) { a[k2 + R2 * k1] = d[k1]; });
    });
}
template <typename T, int SIGN> struct Dft<T, 8, SIGN> {
    static BFPSH_DEV void run(C2<T>* a) { dft_compose<T, 2, 4, SIGN>(a); }
};
template <typename T, int SIGN> struct Dft<T, 16, SIGN> {
    static BFPSH_DEV void run(C2<T>* a) { dft_compose<T, 4, 4, SIGN>(a); }
};
template <typename T, int SIGN> struct Dft<T, 32, SIGN> {
    static BFPSH_DEV void run(C2<T>* a) { dft_compose<T, 4, 8, SIGN>(a); }
};

// ---------------------------------------------------------------- plans
constexpr int ilog2c(int n) {
    int r = 0;
    while ((1 << r) < n) r++;
    return r;
}
struct PlanInfo {
    int ns;
    int r[12];
};
// radices for length N (N = 2^p 3^q): power-of-two part split as evenly as
// possible into stages of radix <= RMAX, then radix-3 stages.
constexpr PlanInfo make_plan(int N, int RMAX) {
    PlanInfo p{};
    p.ns = 0;
    int m = N, n3 = 0;
    while (m % 3 == 0) { m /= 3; n3++; }
    const int bits = ilog2c(m), rb = ilog2c(RMAX);
    const int S = bits == 0 ? 0 : (bits + rb - 1) / rb;
    for (int s = 0; s < S; s++) p.r[p.ns++] = 1 << (bits / S + (s < bits % S ? 1 : 0));
    for (int s = 0; s < n3; s++) p.r[p.ns++] = 3;
    return p;
}
template <int N, int RMAX>
struct Plan {
    static constexpr PlanInfo info = make_plan(N, RMAX);
    static constexpr int nstages = info.ns;
    static constexpr int radix(int s) { return info.r[s]; }
    static constexpr int sublen(int s) {
        int l = 1;
        for (int i = 0; i < s; i++) l *= info.r[i];
        return l;
    }
    static constexpr int max_radix() {
        int m = 1;
        for (int i = 0; i < info.ns; i++) m = info.r[i] > m ? info.r[i] : m;
        return m;
    }
};

// ---------------------------------------------------------------- stages
template <typename T, int N, int B, int LS, int NT, int R, int L, int SIGN>
BFPSH_DEV void stockham_stage(C2<T>* lds, const C2<T>* __restrict__ tw, int tid) {
    constexpr int NB = (N / R) * B;
    constexpr int PER = (NB + NT - 1) / NT;
    C2<T> a[PER][R];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int g = tid + i * NT;
        if (NB % NT == 0 || g < NB) {
            const int b = g % B, j = g / B;
            const int k = j % L;
#pragma unroll
            for (int r = 0; r < R; r++) a[i][r] = lds[(j + r * (N / R)) * LS + b];
            if constexpr (L > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    C2<T> w = tw[r * k * (N / (L * R))];
                    if constexpr (SIGN > 0) w.y = -w.y;
                    a[i][r] = cmul<T>(a[i][r], w);
                }
            }
            Dft<T, R, SIGN>::run(a[i]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int g = tid + i * NT;
        if (NB % NT == 0 || g < NB) {
            const int b = g % B, j = g / B;
            const int k = j % L;
            const int base = (j / L) * L * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) lds[(base + r * L) * LS + b] = a[i][r];
        }
    }
    __syncthreads();
}

// Full FFT of the tile. SIGN = -1 forward, +1 backward (unnormalised).
// Caller must __syncthreads() after filling the tile.
template <typename T, int N, int B, int LS, int NT, int RMAX, int SIGN>
BFPSH_DEV void fft_tile(C2<T>* lds, const C2<T>* __restrict__ tw, int tid) {
    using P = Plan<N, RMAX>;
    sfor<0, P::nstages>([&](auto s) {
        constexpr int R = P::radix(s);
        constexpr int L = P::sublen(s);
        stockham_stage<T, N, B, LS, NT, R, L, SIGN>(lds, tw, tid);
    });
}

}  // namespace bfpsh
