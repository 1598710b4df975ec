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

// cos / sin of 2 pi q / R at compile time (Taylor series in long double on
// the reduced angle; exact to double precision)
constexpr long double kPiL = 3.14159265358979323846264338327950288L;
constexpr long double taylor_sin(long double x) {
    long double term = x, sum = x;
    for (int k = 1; k < 30; k++) {
        term *= -x * x / ((2 * k) * (2 * k + 1));
        sum += term;
    }
    return sum;
}
constexpr long double taylor_cos(long double x) {
    long double term = 1, sum = 1;
    for (int k = 1; k < 30; k++) {
        term *= -x * x / ((2 * k - 1) * (2 * k));
        sum += term;
    }
    return sum;
}
// angle 2 pi q / R with q reduced to [0, R): use the octant symmetry of the
// unit circle so the series argument stays within [0, pi/4]
constexpr double cos_frac(int q, int R) {
    q = ((q % R) + R) % R;
    // 8 q / R in [0, 8): octant index and remainder
    const long double t = (long double)(8 * q) / R;
    const int o = int(t);
    const long double a = (t - o) * kPiL / 4;  // angle within the octant
    const long double b = kPiL / 4 - a;
    switch (o) {
    case 0: return double(taylor_cos(a));
    case 1: return double(taylor_sin(b));
    case 2: return double(-taylor_sin(a));
    case 3: return double(-taylor_cos(b));
    case 4: return double(-taylor_cos(a));
    case 5: return double(-taylor_sin(b));
    case 6: return double(taylor_sin(a));
    default: return double(taylor_cos(b));
    }
}
constexpr double sin_frac(int q, int R) { return cos_frac(4 * q - R, 4 * R); }

// multiply by exp(SIGN 2 pi i M / R), M a compile-time constant
template <typename T, int R, int M, int SIGN>
BFPSH_DEV C2<T> twc(C2<T> a) {
    constexpr int q = ((M % R) + R) % R;
    if constexpr (q == 0) {
        return a;
    } else if constexpr (4 * q == R) {
        return cmul_i<T, SIGN>(a);
    } else if constexpr (2 * q == R) {
        return mk<T>(-a.x, -a.y);
    } else if constexpr (4 * q == 3 * R) {
        return cmul_i<T, -SIGN>(a);
    } else {
        constexpr T c = T(cos_frac(q, R));
        constexpr T s = T(SIGN * sin_frac(q, R));
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
template <typename T, int SIGN> struct Dft<T, 6, SIGN> {
    static BFPSH_DEV void run(C2<T>* a) { dft_compose<T, 2, 3, SIGN>(a); }
};
template <typename T, int SIGN> struct Dft<T, 12, SIGN> {
    static BFPSH_DEV void run(C2<T>* a) { dft_compose<T, 4, 3, SIGN>(a); }
};
template <typename T, int SIGN> struct Dft<T, 24, SIGN> {
    static BFPSH_DEV void run(C2<T>* a) { dft_compose<T, 8, 3, SIGN>(a); }
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
// radices for the register-resident FFT: every radix divides E (each thread
// then owns whole butterflies at every stage), largest first
constexpr PlanInfo make_plan_e(int N, int E) {
    PlanInfo p{};
    p.ns = 0;
    const int allowed[9] = {32, 24, 16, 12, 8, 6, 4, 3, 2};
    int m = N;
    while (m > 1 && p.ns < 12) {
        int pick = 0;
        for (int i = 0; i < 9; i++)
            if (E % allowed[i] == 0 && m % allowed[i] == 0) { pick = allowed[i]; break; }
        if (pick == 0) { p.ns = -1; return p; }
        p.r[p.ns++] = pick;
        m /= pick;
    }
    return p;
}
template <int N, int E>
struct PlanE {
    static constexpr PlanInfo info = make_plan_e(N, E);
    static_assert(info.ns > 0, "no radix plan for this length / register count");
    static constexpr int nstages = info.ns;
    static constexpr int radix(int s) { return info.r[s]; }
    static constexpr int sublen(int s) {
        int l = 1;
        for (int i = 0; i < s; i++) l *= info.r[i];
        return l;
    }
};

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

namespace bfpsh {

// ---------------------------------------------------------------- register-resident Stockham
// Tile of B columns x N rows, NT threads: column b = t % B, group g = t / B
// (G = NT / B groups per column).  Every thread holds E = N / G complex
// values of its column in registers.  At stage s (radix R, sub-length L) a
// thread owns butterflies j = g + i G, i < N / (R G), and its register
// element e = i R + r is butterfly input r, i.e. row j + r N / R.  After the
// stage the element holds output r at row (j / L) L R + j mod L + r L; the
// rows are exchanged through the LDS tile (element (n, b) at lds[n LS + b])
// between stages only.  Callers load stage-0 inputs (rows in_row(g, e)) and
// read last-stage outputs (rows out_row(g, e)) directly from registers.
template <typename T, int N, int B, int LS, int NT, int SIGN>
struct RegFft {
    static constexpr int G = NT / B;
    static constexpr int E = N / G;
    using P = PlanE<N, E>;
    static constexpr int S = P::nstages;
    static_assert(NT % B == 0 && N % G == 0, "tile shape");

    static constexpr int per(int s) { return N / (P::radix(s) * G); }

    // row of register element e at the input of stage s
    static BFPSH_DEV int in_row_s(int s, int g, int e) {
        const int R = P::radix(s);
        const int i = e / R, r = e % R;
        return g + i * G + r * (N / R);
    }
    static BFPSH_DEV int out_row_s(int s, int g, int e) {
        const int R = P::radix(s), L = P::sublen(s);
        const int i = e / R, r = e % R;
        const int j = g + i * G;
        return (j / L) * L * R + j % L + r * L;
    }
    static BFPSH_DEV int in_row(int g, int e) { return in_row_s(0, g, e); }
    static BFPSH_DEV int out_row(int g, int e) { return out_row_s(S - 1, g, e); }

    template <int s>
    static BFPSH_DEV void compute_stage(C2<T> (&v)[E], const C2<T>* __restrict__ tw, int g) {
        constexpr int R = P::radix(s), L = P::sublen(s), PER = N / (R * G);
        static_assert(PER >= 1, "radix larger than the register set");
#pragma unroll
        for (int i = 0; i < PER; i++) {
            C2<T>* a = v + i * R;
            if constexpr (L > 1) {
                const int k = (g + i * G) % L;
#pragma unroll
                for (int r = 1; r < R; r++) {
                    C2<T> w = tw[r * k * (N / (L * R))];
                    if constexpr (SIGN > 0) w.y = -w.y;
                    a[r] = cmul<T>(a[r], w);
                }
            }
            Dft<T, R, SIGN>::run(a);
        }
    }

    // v: stage-0 inputs on entry, last-stage outputs on exit.  The LDS tile
    // is free again when this returns (trailing barrier included only when
    // an exchange happened).
    static BFPSH_DEV void run(C2<T> (&v)[E], C2<T>* lds, const C2<T>* __restrict__ tw, int g, int b) {
        sfor<0, S>([&](auto s) {
            compute_stage<s>(v, tw, g);
            if constexpr (s + 1 < S) {
#pragma unroll
                for (int e = 0; e < E; e++) lds[out_row_s(s, g, e) * LS + b] = v[e];
                __syncthreads();
#pragma unroll
                for (int e = 0; e < E; e++) v[e] = lds[in_row_s(s + 1, g, e) * LS + b];
                __syncthreads();
            }
        });
    }
};

}  // namespace bfpsh
