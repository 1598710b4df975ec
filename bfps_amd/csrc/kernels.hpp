// HIP kernels of the vorticity-equation step and the tracer loop (gfx950).
//
// Device data layout (one GPU holds the whole periodic box, cubic N^3):
//   spectral plane  [ky][kz][kx],  kx = 0 .. N/2-1 stored (NXH = N/2)
//   real plane      [y][z][x],     x  = 0 .. N-1
// A vector field is 3 (or 6, 9) planes plane_size(N) complex values apart;
// within a plane, ky (or y) rows are row_stride(N) = N*NXH + BFPSH_PAD_Y
// complex apart (common.hpp), and a real row (y, z) occupies the bytes of
// the spectral row (ky = y, kz = z).  The kx = N/2 plane of the reference's
// half spectrum is not stored: it lies outside the dealiasing sphere
// (kx^2 >= (N/2)^2 > kM^2), the reference's io_checkpoint read zeroes it
// (low_pass, bfps/cpp/kspace.cpp:235-253) and the time step never writes it
// (bfps/cpp/vorticity_equation.cpp:372), so it is identically zero.
//
// Inverse 3D transform = z pass (within ky slab) -> y pass -> x (c2r) pass;
// forward = x (r2c) -> z -> y.  The real-space array therefore comes out
// as [y][z][x]; the x passes work on pairs of rows (z, z+1) packed as the
// real and imaginary parts of one complex sequence.
#pragma once
#include "fft.hpp"

#ifndef BFPSH_TILE_BYTES
#define BFPSH_TILE_BYTES 128
#endif
// LDS bytes a tile of the plain strided pass may take (B columns x N
// complex values; gfx950 has 160 KB of LDS per CU)
#ifndef BFPSH_TILE_LDS
#define BFPSH_TILE_LDS 131072
#endif
// register-resident x pass (k_rows_cross_w): minimum waves per SIMD and
// LDS padding (one slot every 2^BFPSH_RW_PAD rows); measured at 512^3 FP32,
// 3 waves with padding every 16 rows beat 2 waves with padding every 8
#ifndef BFPSH_RW_WAVES
#define BFPSH_RW_WAVES 3
#endif
#ifndef BFPSH_RW_PAD
#define BFPSH_RW_PAD 4
#endif
// 1: x-axis rows through the LDS-tile FFT for every size (A/B switch)
#ifndef BFPSH_ROWS_TILE
#define BFPSH_ROWS_TILE 0
#endif

namespace bfpsh {

// ------------------------------------------------------------ configuration
template <typename T, int N>
struct Cfg {
    static constexpr int NXH = N / 2;
    static constexpr int RMAX = 16;  // radix cap of the LDS-tile FFT (row passes)
    // strided passes: register-resident FFT, E complex values per thread and
    // field (all radices divide E)
    static constexpr int ereg() {
        if (N % 3 == 0) return 12;
        if (N <= 16) return N;
        if (sizeof(T) == 4 && (N == 128 || N == 256 || N == 1024)) return 16;
        return 8;
    }
    static constexpr int E = ereg();
    static constexpr int G = N / E;  // thread groups per column
    // columns per strided tile: BFPSH_TILE_BYTES-byte row segments (default
    // 128) in at most lds bytes of LDS
    static constexpr int bcap(int lds) {
        int b = int(BFPSH_TILE_BYTES / sizeof(C2<T>));
        const int lim = int(lds / (N * sizeof(C2<T>)));
        if (b > lim) b = lim;
        if (b > NXH) b = NXH;
        while (NXH % b) b /= 2;
        return b;
    }
    // plain strided pass (k_pass): up to BFPSH_TILE_LDS bytes, which only
    // widens the 1024^3 tiles (FP32 8 -> 16 columns: one y pass over a plane
    // 2.97 -> 2.46 ms); the fused first and last passes keep 64 KB tiles
    // (their three components in registers need the smaller workgroups)
    static constexpr int B = bcap(BFPSH_TILE_LDS);
    static constexpr int BF = bcap(65536);
    static constexpr int NT = G * B;
    // Biot-Savart pass: three source components stay in registers; half the
    // columns (when NT would exceed 512) keep it within a 256-VGPR budget
#ifndef BFPSH_E_BS
#define BFPSH_E_BS 0
#endif
    static constexpr int E_BS = (sizeof(T) == 4 && N == 512 && BFPSH_E_BS > 0 && N % BFPSH_E_BS == 0) ? BFPSH_E_BS : E;
    static constexpr int G_BS = N / E_BS;
    static constexpr int B_BS = (G_BS * BF > 512 && BF > 1) ? BF / 2 : BF;
    static constexpr int NT_BS = G_BS * B_BS;
    // last forward pass with the update epilogue (3 components in registers)
    static constexpr int B_Y3 = BF;
    // elements per thread of the plain strided pass (measured at 512^3 FP32:
    // 16 elements x 512 threads beat 8 x 1024) and of the last forward pass
    // (E; the register budget of its epilogue favours more threads)
#ifndef BFPSH_E_PASS
#define BFPSH_E_PASS 16
#endif
#ifndef BFPSH_E_Y3
#define BFPSH_E_Y3 0
#endif
    static constexpr int pick_e(int e, int b) { return (e > 0 && N % e == 0 && (N / e) * b <= 1024) ? e : E; }
    static constexpr int E_PASS = pick_e(sizeof(T) == 4 && N == 512 ? BFPSH_E_PASS : 0, B);
    static constexpr int NT_PASS = (N / E_PASS) * B;
    static constexpr int E_Y3 = pick_e(sizeof(T) == 4 && N == 512 ? BFPSH_E_Y3 : 0, B_Y3);
    static constexpr int NT_Y3 = (N / E_Y3) * B_Y3;
    static constexpr int RP = (N * int(sizeof(C2<T>)) >= 8192) ? 1 : 2;  // row pairs per row block
    // threads of the LDS-tile row passes (1024: two blocks of 512 per CU,
    // four waves per SIMD, instead of two blocks of 256)
    static constexpr int NT_ROW = N >= 1024 ? 512 : (N >= 256 ? 256 : (N >= 64 ? 128 : 64));
    // LDS row stride of the LDS-tile cross-product pass (6 RP components per
    // row): odd, spreading strided lanes over the banks, unless that slot
    // costs a resident workgroup per CU (1024^3 FP32: three 48 KB
    // workgroups instead of two of 56 KB, 14.4 -> 12.0 ms per launch)
    static constexpr int ls_row() {
        const long c = 6 * RP, lds = 160 * 1024, row = long(N) * long(sizeof(C2<T>));
        return lds / (row * c) > lds / (row * (c + 1)) ? int(c) : int(c + 1);
    }
    static constexpr int LS_ROW = ls_row();
    static_assert(NT <= 1024, "tile threads");
};

// offsets of row (y, z), row = y N + z, of a plane: complex values of a
// spectral row and reals of a real row (the same bytes)
template <int N> BFPSH_DEV size_t crow(size_t row) { return (row / N) * row_stride(N) + (row % N) * (N / 2); }
template <int N> BFPSH_DEV size_t rrow(size_t row) { return (row / N) * 2 * row_stride(N) + (row % N) * N; }

// ------------------------------------------------------------ parameters
struct Geom {
    int n;
    double dkx, dky, dkz;
    double kM2;
    int smooth;  // 1: smooth filter exp(-36 (k^2/kM^2)^18), 0: two-thirds low pass
};
struct StepParams {
    double nu, dt, famp, fk0, fk1;
    int forcing;  // 0 none, 1 linear, 2 Kolmogorov, 3 fixed energy-injection rate
    int fmode;
    const double* famp_dev;  // forcing 3: band amplitudes of the 3 stages (device)
};

BFPSH_DEV double kwave(int i, int n, double dk) { return (i <= n / 2 ? i : i - n) * dk; }

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2).
// Remap a 2D grid so that each XCD walks a contiguous range of logical blocks
// (x fastest): neighbouring x tiles, which share 128-byte lines when a tile
// row is 64 bytes, then run on the same XCD back to back.
BFPSH_DEV void xcd_tile(int& bx, int& by) {
    const unsigned gx = gridDim.x, total = gridDim.x * gridDim.y;
    unsigned p = blockIdx.x + blockIdx.y * gx;
    if (total % 8u == 0) p = (p % 8u) * (total / 8u) + p / 8u;
    bx = int(p % gx);
    by = int(p / gx);
}

// ------------------------------------------------------------ K1 + first inverse pass
// MODE 0: outputs u_hat (3), omega_hat (3)        [nonlinear term]
// MODE 1: outputs u_hat (3)                        [velocity in real space]
// MODE 2: outputs i k_j u_hat_i (9), k^2 < kM^2    [velocity gradient]
// u_hat = i k x omega_hat / k^2 for 0 < k^2 <= kM^2, else 0
// (compute_velocity, bfps/cpp/vorticity_equation.cpp:206-239, with the
// symmetrize of field.cpp:900-978 applied to its kx = 0 plane by reading the
// conjugated mirror mode; compute_gradient, bfps/cpp/field.cpp:1077-1097).
template <int MODE> struct BsOut { static constexpr int NF = MODE == 0 ? 6 : (MODE == 1 ? 3 : 9); };

// MIRROR: the launch covers the kx tile holding kx = 0 and applies the
// symmetrize substitution there; the other tiles use MIRROR = false
// (blockIdx.x offset by tile0) and carry no mirror code.
template <typename T, int N, int MODE, bool MIRROR, int E_ = Cfg<T, N>::E_BS, int B_ = Cfg<T, N>::B_BS>
__global__ void __launch_bounds__(((N / E_) * B_))
k_inv_z_bs(const C2<T>* __restrict__ w, C2<T>* __restrict__ out, const C2<T>* __restrict__ tw, Geom g, int tile0,
           size_t bstride) {
    using CF = Cfg<T, N>;
    // box blockIdx.z of a batch of equal boxes bstride elements apart
    w += blockIdx.z * bstride;
    out += blockIdx.z * bstride;
    constexpr int B = B_, E = E_, NT = (N / E_) * B_, NXH = CF::NXH;
    constexpr size_t PS = plane_size(N), YS = row_stride(N);
    constexpr int NF = BsOut<MODE>::NF;
    using F = RegFft<T, N, B, B, NT, +1>;
    __shared__ C2<T> lds[N * B];
    const int tid = threadIdx.x;
    const int b = tid % B, grp = tid / B;
    int tile, ky;
    xcd_tile(tile, ky);
    const int kxi = (tile + tile0) * B + b;
    const double kx = kxi * g.dkx, kyv = kwave(ky, N, g.dky);
    C2<T> wr[3][E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int n = F::in_row(grp, e);
        const size_t off = size_t(ky) * YS + size_t(n) * NXH + kxi;
#pragma unroll
        for (int c = 0; c < 3; c++) wr[c][e] = w[c * PS + off];
    }
    // Velocity-type fields in a runtime loop (register arrays indexed by
    // constants only), then the vorticity copies (MODE 0) as straight-line
    // code: a runtime select between wr[0..2] would turn into a pointer into
    // wr and force the array into scratch memory.
    constexpr int NU = MODE == 0 ? 3 : NF;
#pragma unroll 1
    for (int f = 0; f < NU; f++) {
        // an opaque copy of the group index: the per-element wavenumbers are
        // recomputed per output field instead of being held across the loop
        int gq = grp;
        asm volatile("" : "+v"(gq));
        // likewise the source values: products of them with the wavenumbers
        // would otherwise be hoisted (in FP64) out of the loop
#pragma unroll
        for (int e = 0; e < E; e++)
#pragma unroll
            for (int c = 0; c < 3; c++) asm volatile("" : "+v"(wr[c][e].x), "+v"(wr[c][e].y));
        C2<T> v[E];
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int n = F::in_row(gq, e);
            const double kz = kwave(n, N, g.dkz);
            const double k2 = kx * kx + kyv * kyv + kz * kz;
            const int cu = MODE == 2 ? f % 3 : f;
            C2<T> a0 = wr[0][e], a1 = wr[1][e], a2 = wr[2][e];
            // symmetrized kx = 0 plane: the lower half takes conj(mirror)
            if (MIRROR && kxi == 0 && ((ky > N / 2) || (ky == 0 && n > N / 2))) {
                const size_t moff = size_t((N - ky) % N) * YS + size_t((N - n) % N) * NXH;
                a0 = cconj<T>(w[moff]);
                a1 = cconj<T>(w[PS + moff]);
                a2 = cconj<T>(w[2 * PS + moff]);
            }
            C2<T> u = mk<T>(T(0), T(0));
            if (k2 <= g.kM2 && k2 > 0) {
                // one FP64 division per element and field, then products
                // (within an ulp of the quotients before the rounding to T)
                const double ik2 = 1.0 / k2;
                if (cu == 0) u = mk<T>(T(-(kyv * a2.y - kz * a1.y) * ik2), T((kyv * a2.x - kz * a1.x) * ik2));
                if (cu == 1) u = mk<T>(T(-(kz * a0.y - kx * a2.y) * ik2), T((kz * a0.x - kx * a2.x) * ik2));
                if (cu == 2) u = mk<T>(T(-(kx * a1.y - kyv * a0.y) * ik2), T((kx * a1.x - kyv * a0.x) * ik2));
            }
            if (MODE == 2) {
                const int j = f / 3;
                const double kj = j == 0 ? kx : (j == 1 ? kyv : kz);
                v[e] = k2 < g.kM2 ? mk<T>(T(-kj * u.y), T(kj * u.x)) : mk<T>(T(0), T(0));
            } else {
                v[e] = u;
            }
        }
        F::run(v, lds, tw, gq, b);
        C2<T>* o = out + size_t(f) * PS;
#pragma unroll
        for (int e = 0; e < E; e++) o[size_t(ky) * YS + size_t(F::out_row(gq, e)) * NXH + kxi] = v[e];
    }
    if constexpr (MODE == 0) {
        sfor<0, 3>([&](auto cc) {
            constexpr int c = cc;
            F::run(wr[c], lds, tw, grp, b);
            C2<T>* o = out + size_t(3 + c) * PS;
#pragma unroll
            for (int e = 0; e < E; e++) o[size_t(ky) * YS + size_t(F::out_row(grp, e)) * NXH + kxi] = wr[c][e];
        });
    }
}

// ------------------------------------------------------------ generic strided pass (in place)
// element (n, b) of tile (blockIdx.x, blockIdx.y) of plane p of box q
// (blockIdx.z = q * planes + p) at
//   base + q * bstride + p * plane_size(N) + blockIdx.y * ostride + n * estride + kx0 + b
// Tile shape: B columns, E complex values per thread (N / E threads per
// column); the shapes a context may use are listed by pass_shapes() and
// picked by the pass autotuner.
template <typename T, int N, int SIGN, int E, int B>
__global__ void __launch_bounds__((N / E) * B)
k_pass(C2<T>* __restrict__ data, size_t estride, size_t ostride, const C2<T>* __restrict__ tw, int planes,
       size_t bstride) {
    constexpr int NT = (N / E) * B;
    using F = RegFft<T, N, B, B, NT, SIGN>;
    static_assert(F::E == E, "tile shape");
    constexpr size_t PS = plane_size(N);
    __shared__ C2<T> lds[N * B];
    const int tid = threadIdx.x;
    const int b = tid % B, grp = tid / B;
    // blockIdx.z = box * planes + plane (boxes bstride elements apart)
    const int box = blockIdx.z / planes, plane = blockIdx.z % planes;
    C2<T>* base = data + box * bstride + plane * PS + blockIdx.y * ostride + blockIdx.x * B + b;
    C2<T> v[E];
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = base[F::in_row(grp, e) * estride];
    F::run(v, lds, tw, grp, b);
#pragma unroll
    for (int e = 0; e < E; e++) base[F::out_row(grp, e) * estride] = v[e];
}

// Persistent form of k_pass: gridDim.x workgroups (as many as fit on the
// chip at once) each walk tiles t = blockIdx.x + k gridDim.x of the
// ntx x nty x (planes x boxes) tile grid, and load the columns of their
// next tile while the current one is transformed and stored, so that
// global-memory latency hides behind the FFT (the barriers of the FFT
// order LDS only and leave those loads in flight).
template <typename T, int N, int SIGN, int E, int B>
__global__ void __launch_bounds__((N / E) * B)
k_pass_pp(C2<T>* __restrict__ data, size_t estride, size_t ostride, const C2<T>* __restrict__ tw, int planes,
          size_t bstride, int ntx, int nty, int ntiles) {
    constexpr int NT = (N / E) * B;
    using F = RegFft<T, N, B, B, NT, SIGN>;
    static_assert(F::E == E, "tile shape");
    constexpr size_t PS = plane_size(N);
    __shared__ C2<T> lds[N * B];
    const int tid = threadIdx.x;
    const int b = tid % B, grp = tid / B;
    auto base_of = [&](int t) {
        const int x = t % ntx, r = t / ntx, y = r % nty, zp = r / nty;
        const int box = zp / planes, plane = zp % planes;
        return data + box * bstride + plane * PS + size_t(y) * ostride + size_t(x) * B + b;
    };
    int t = blockIdx.x;
    if (t >= ntiles) return;
    C2<T> cur[E], nxt[E];
    {
        const C2<T>* p = base_of(t);
#pragma unroll
        for (int e = 0; e < E; e++) cur[e] = p[F::in_row(grp, e) * estride];
    }
    for (;;) {
        const int tn = t + int(gridDim.x);
        const bool more = tn < ntiles;  // uniform
        if (more) {
            const C2<T>* p = base_of(tn);
#pragma unroll
            for (int e = 0; e < E; e++) nxt[e] = p[F::in_row(grp, e) * estride];
        }
        F::run(cur, lds, tw, grp, b);
        {
            C2<T>* p = base_of(t);
#pragma unroll
            for (int e = 0; e < E; e++) p[F::out_row(grp, e) * estride] = cur[e];
        }
        if (!more) break;
#pragma unroll
        for (int e = 0; e < E; e++) cur[e] = nxt[e];
        t = tn;
    }
}

// Tile shapes (E values per thread, B columns, pp: the persistent
// prefetching form) of the plain strided pass for one (precision, N): the
// built-in choice (Cfg) first, then, for N >= 256 (below that a box is
// cache-resident and launch-bound), every shape with E in {8, 12, 16, 24,
// 32} dividing N with a radix plan, B dividing N/2, 64..1024 threads, at
// least 32-byte row segments and at most 128 KB of LDS, and the persistent
// form of each whose two register tiles (4 E registers of FP32, 8 E of
// FP64) take at most 2/5 of a lane's registers (beyond that it spills) --
// up to kMaxPassShapes in all.
struct PassShape {
    int e, b, pp;
};
constexpr int kMaxPassShapes = 16;
struct PassShapes {
    int n;
    PassShape v[kMaxPassShapes];
};
template <typename T, int N>
constexpr PassShapes pass_shapes() {
    using CF = Cfg<T, N>;
    PassShapes s{};
    s.n = 0;
    s.v[s.n++] = PassShape{CF::E_PASS, CF::B, 0};
    if (N < 256) return s;
    constexpr int es[5] = {8, 12, 16, 24, 32};
    constexpr int bs[6] = {2, 4, 8, 16, 32, 64};
    for (int pp = 0; pp < 2; pp++)
        for (int i = 0; i < 5; i++) {
            const int e = es[i];
            if (N % e != 0 || make_plan_e(N, e).ns <= 0) continue;
            for (int j = 0; j < 6; j++) {
                const int b = bs[j];
                const long nt = long(N / e) * b;
                if ((N / 2) % b != 0 || nt < 64 || nt > 1024) continue;
                // persistent form: its two register tiles at most 2/5 of the
                // registers a lane gets (256, or 128 in 1024-thread workgroups)
                if (pp && 5 * (2 * e * int(sizeof(C2<T>)) / 4) > 2 * (nt > 512 ? 128 : 256)) continue;
                if (long(N) * b * long(sizeof(C2<T>)) > 131072 || b * int(sizeof(C2<T>)) < 32) continue;
                bool dup = false;
                for (int k = 0; k < s.n; k++) dup |= s.v[k].e == e && s.v[k].b == b && s.v[k].pp == pp;
                if (dup || s.n == kMaxPassShapes) continue;
                s.v[s.n++] = PassShape{e, b, pp};
            }
        }
    return s;
}

// Tile shapes of the fused first (Biot-Savart + inverse z) and last
// (forward y + update) passes, which hold three components in registers:
// the built-in choice first, then for N >= 512 the shapes with E in
// {8, 16}, 64..1024 threads, at least 64-byte row segments, at most 128 KB
// of LDS, and three components of E values that take at most 96 registers
// of a lane (48 in 1024-thread workgroups, whose lanes get 128 registers):
// larger shapes spill.
template <typename T, int N>
constexpr PassShapes fused_shapes(int e0, int b0) {
    PassShapes s{};
    s.n = 0;
    s.v[s.n++] = PassShape{e0, b0, 0};
    if (N < 512) return s;
    constexpr int es[2] = {8, 16};
    constexpr int bs[4] = {4, 8, 16, 32};
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 4; j++) {
            const int e = es[i], b = bs[j];
            const long nt = long(N / e) * b;
            if (N % e != 0 || (N / 2) % b != 0 || nt < 64 || nt > 1024) continue;
            if (long(N) * b * long(sizeof(C2<T>)) > 131072 || b * int(sizeof(C2<T>)) < 64) continue;
            const int regs = 3 * e * int(sizeof(C2<T>)) / 4;  // VGPRs of the three components
            if (regs > (nt > 512 ? 48 : 96)) continue;  // beyond these the compiler spills
            bool dup = false;
            for (int k = 0; k < s.n; k++) dup |= s.v[k].e == e && s.v[k].b == b;
            if (dup || s.n == kMaxPassShapes) continue;
            s.v[s.n++] = PassShape{e, b, 0};
        }
    return s;
}

// ------------------------------------------------------------ x passes on row pairs
// Row pair q of row block: rows (y, z = 2q) and (y, z = 2q+1) of a plane.
// Inverse: Z[k] = A[k] + i B[k] (k < N/2; k = 0 real parts only, as a c2r
// ignores the imaginary part of the zero mode), Z[N/2] = 0,
// Z[N-k] = conj(A[k]) + i conj(B[k]); z = IDFT(Z); row z = Re z, row z+1 = Im z.
// Forward: Z = DFT(a + i b); A[k] = (Z[k] + conj Z[N-k]) / 2,
// B[k] = (Z[k] - conj Z[N-k]) / (2 i).
// NCOL columns at once: all global loads of a thread are issued before its
// LDS stores.  Column col reads rows ra(col), rb(col) = base + off[col].
template <typename T, int N, int LS, int NCOL, int NT>
BFPSH_DEV void rows_unpack_inverse(C2<T>* lds, const C2<T>* const (&ra)[NCOL], const C2<T>* const (&rb)[NCOL],
                                   int tid) {
    constexpr int H = N / 2;
    constexpr int PER = (H + NT - 1) / NT;
#pragma unroll
    for (int it = 0; it < PER; it++) {
        const int k = tid + it * NT;
        if (H % NT != 0 && k >= H) break;
        C2<T> A[NCOL], Bv[NCOL];
#pragma unroll
        for (int col = 0; col < NCOL; col++) {
            A[col] = ra[col][k];
            Bv[col] = rb[col][k];
        }
#pragma unroll
        for (int col = 0; col < NCOL; col++) {
            if (k == 0) {
                lds[col] = mk<T>(A[col].x, Bv[col].x);
                lds[H * LS + col] = mk<T>(T(0), T(0));
            } else {
                lds[k * LS + col] = mk<T>(A[col].x - Bv[col].y, A[col].y + Bv[col].x);
                lds[(N - k) * LS + col] = mk<T>(A[col].x + Bv[col].y, Bv[col].x - A[col].y);
            }
        }
    }
}
template <typename T, int N, int LS>
BFPSH_DEV void rows_pack_forward(const C2<T>* lds, int col, C2<T>* ra, C2<T>* rb, int tid, int nt) {
    constexpr int H = N / 2;
    for (int k = tid; k < H; k += nt) {
        const C2<T> Z = lds[k * LS + col];
        const C2<T> M = lds[((N - k) % N) * LS + col];
        ra[k] = mk<T>(T(0.5) * (Z.x + M.x), T(0.5) * (Z.y - M.y));
        rb[k] = mk<T>(T(0.5) * (Z.y + M.y), T(-0.5) * (Z.x - M.x));
    }
}

// u x omega / N^3 in real space between the inverse and forward x passes
// (omega_nonlin, bfps/cpp/vorticity_equation.cpp:297-323).
// in:  planes 0..2 = u_hat, 3..5 = omega_hat after the z and y inverse passes
// out: planes 0..2 = x-transformed (u x omega)/N^3 (in place); optional real
//      velocity ru (3 real planes [y][z][x]) for the tracers.
template <typename T, int N>
__global__ void __launch_bounds__((Cfg<T, N>::NT_ROW))
k_rows_cross(C2<T>* __restrict__ data, T* __restrict__ ru, const C2<T>* __restrict__ tw, T npoints, size_t bstride) {
    using CF = Cfg<T, N>;
    // LDS row stride: odd, so lanes striding by rows hit distinct banks
    constexpr int RP = CF::RP, NT = CF::NT_ROW, NXH = CF::NXH, LS = CF::LS_ROW;
    constexpr size_t PS = plane_size(N), YS = row_stride(N);
    __shared__ C2<T> lds[N * LS];
    // box blockIdx.y of a batch (complex planes and real velocity move together)
    data += blockIdx.y * bstride;
    if (ru) ru += blockIdx.y * 2 * bstride;
    // the 1/npoints normalization as a product (exact for the power-of-two
    // sizes, within an ulp of the quotient otherwise)
    const T inp = T(1) / npoints;
    const int tid = threadIdx.x;
    const int q0 = blockIdx.x * RP;  // first row pair; rows (y, z) flattened: row = y*N + z
    {
        const C2<T>* ra[6 * RP];
        const C2<T>* rb[6 * RP];
#pragma unroll
        for (int c = 0; c < 6; c++)
#pragma unroll
            for (int p = 0; p < RP; p++) {
                const size_t row = size_t(q0 + p) * 2;
                ra[c * RP + p] = data + c * PS + crow<N>(row);
                rb[c * RP + p] = data + c * PS + crow<N>(row) + NXH;
            }
        rows_unpack_inverse<T, N, LS, 6 * RP, NT>(lds, ra, rb, tid);
    }
    __syncthreads();
    fft_tile<T, N, 6 * RP, LS, NT, CF::RMAX, +1>(lds, tw, tid);
    for (int e = tid; e < N * RP; e += NT) {
        const int nn = e / RP, p = e % RP;
        C2<T> u[3], w[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            u[c] = lds[nn * LS + c * RP + p];
            w[c] = lds[nn * LS + (3 + c) * RP + p];
        }
        if (ru) {
            const size_t row = size_t(q0 + p) * 2;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                ru[c * 2 * PS + rrow<N>(row) + nn] = u[c].x;
                ru[c * 2 * PS + rrow<N>(row) + N + nn] = u[c].y;
            }
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            const T re = u[c1].x * w[c2].x - u[c2].x * w[c1].x;
            const T im = u[c1].y * w[c2].y - u[c2].y * w[c1].y;
            lds[nn * LS + c * RP + p] = mk<T>(re * inp, im * inp);
        }
    }
    __syncthreads();
    fft_tile<T, N, 3 * RP, LS, NT, CF::RMAX, -1>(lds, tw, tid);
    for (int c = 0; c < 3; c++)
        for (int p = 0; p < RP; p++) {
            const size_t row = size_t(q0 + p) * 2;
            rows_pack_forward<T, N, LS>(lds, c * RP + p, data + c * PS + crow<N>(row),
                                        data + c * PS + crow<N>(row) + NXH, tid, NT);
        }
}

// Register-resident form of k_rows_cross (FP32, N <= 512): G = N / 8 threads
// own one row pair and hold 8 values of each of the six components, so the
// cross product is formed in registers; the radix exchanges of each
// component go through an LDS region private to the row pair (with padding
// slots against bank conflicts), and the pairing of Z[k] with Z[N-k] in the packed
// rows is a lane shuffle.  A row pair lives inside one wave (G <= 64), so the
// exchanges need wave-level ordering only, no workgroup barrier.  Element e
// of a thread holds row g + G sigma(e) at the input of the FFT and row
// g + G tau(e) at its output (g = thread index within the row pair).
template <typename T, int N>
struct RowsW {
    static constexpr int E = 8, G = N / E, NT = 256, RPB = NT / G, H = E / 2;
    static constexpr int PAD = BFPSH_RW_PAD;  // one LDS slot of padding every 2^PAD rows
    static constexpr int SL = N + (N >> PAD);  // LDS slots per row pair
    static constexpr bool ok = sizeof(T) == 4 && N >= 64 && N <= 512 && (N & (N - 1)) == 0;
    using P = PlanE<N, E>;
    static constexpr int S = P::nstages;
    static constexpr int sigma(int e) {
        const int R = P::radix(0);
        return e / R + (e % R) * (E / R);
    }
    static constexpr int tau(int e) {
        const int R = P::radix(S - 1);
        return e / R + (e % R) * (E / R);
    }
    static constexpr int tau_inv(int m) {
        for (int e = 0; e < E; e++)
            if (tau(e) == m) return e;
        return -1;
    }
    static BFPSH_DEV int slot(int row) { return row + (row >> PAD); }
};

BFPSH_DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// FFT of NCG components of a row pair at once (their exchanges share the
// wave-level ordering points): stage inputs in v, outputs in v
template <typename T, int N, int SIGN, int NCG>
BFPSH_DEV void rows_w_fft(C2<T> (&v)[NCG][8], C2<T>* reg, const C2<T>* __restrict__ tw, int g) {
    using RW = RowsW<T, N>;
    using F = RegFft<T, N, 1, 1, RW::G, SIGN>;
    // an opaque copy of the twiddle pointer: the twiddles are re-read (from
    // cache) per group instead of being held in registers across all of them
    const C2<T>* twl = tw;
    asm volatile("" : "+s"(twl));
    // likewise the LDS slot addresses are recomputed per group
    int gl = g;
    asm volatile("" : "+v"(gl));
    sfor<0, RW::S>([&](auto s) {
#pragma unroll
        for (int c = 0; c < NCG; c++) F::template compute_stage<s>(v[c], twl, gl);
        if constexpr (s + 1 < RW::S) {
#pragma unroll
            for (int c = 0; c < NCG; c++)
#pragma unroll
                for (int e = 0; e < 8; e++) reg[c * RW::SL + RW::slot(F::out_row_s(s, gl, e))] = v[c][e];
            wave_sync();
#pragma unroll
            for (int c = 0; c < NCG; c++)
#pragma unroll
                for (int e = 0; e < 8; e++) v[c][e] = reg[c * RW::SL + RW::slot(F::in_row_s(s + 1, gl, e))];
            wave_sync();
        }
    });
}

template <typename T>
BFPSH_DEV C2<T> shfl2(C2<T> a, int src) {
    return mk<T>(__shfl(a.x, src, 64), __shfl(a.y, src, 64));
}

template <typename T, int N>
__global__ void __launch_bounds__(256, BFPSH_RW_WAVES)
k_rows_cross_w(C2<T>* __restrict__ data, T* __restrict__ ru, const C2<T>* __restrict__ tw, T npoints,
               size_t bstride) {
    using RW = RowsW<T, N>;
    constexpr int E = RW::E, G = RW::G, H = RW::H, RPB = RW::RPB, NXH = N / 2;
    constexpr size_t PS = plane_size(N), YS = row_stride(N);
    static_assert(RW::ok && 64 % G == 0, "row pair inside one wave");
    __shared__ C2<T> lds[RPB * 3 * RW::SL];
    data += blockIdx.y * bstride;
    if (ru) ru += blockIdx.y * 2 * bstride;
    const T inp = T(1) / npoints;  // normalization as a product, as k_rows_cross
    const int tid = threadIdx.x;
    const int rp = tid / G, g = tid % G;
    const int partner = (tid & 63) - g + (G - g) % G;  // lane holding index N - k
    C2<T>* reg = lds + rp * 3 * RW::SL;
    const size_t q = size_t(blockIdx.x) * RPB + rp;  // rows 2q, 2q + 1 (row = y N + z)
    const size_t ra = crow<N>(2 * q), rb = ra + NXH;
    C2<T> v[6][E];
    // unpack (as rows_unpack_inverse): Z[k] = A[k] + i B[k], Z[N/2] = 0,
    // Z[N-k] = conj(A[k]) + i conj(B[k]); rows k = g + G m, m < E / 2.  The
    // loads of omega are issued before the FFTs of u.
    C2<T> A[3][H], B[3][H];
    auto load = [&](int c0) {
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int m = 0; m < H; m++) {
                A[c][m] = data[(c0 + c) * PS + ra + g + G * m];
                B[c][m] = data[(c0 + c) * PS + rb + g + G * m];
            }
    };
    auto unpack = [&](C2<T> (&vv)[3][E]) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            C2<T> z[E];
#pragma unroll
            for (int m = 0; m < H; m++) {
                z[m] = mk<T>(A[c][m].x - B[c][m].y, A[c][m].y + B[c][m].x);
                if (m == 0 && g == 0) z[m] = mk<T>(A[c][m].x, B[c][m].x);
            }
#pragma unroll
            for (int m = H; m < E; m++) {
                C2<T> a = shfl2<T>(A[c][E - 1 - m], partner), b = shfl2<T>(B[c][E - 1 - m], partner);
                if (g == 0) {
                    if (m == H) {
                        a = mk<T>(T(0), T(0));
                        b = a;
                    } else {
                        a = A[c][(E - m) % H];
                        b = B[c][(E - m) % H];
                    }
                }
                z[m] = mk<T>(a.x + b.y, b.x - a.y);
            }
#pragma unroll
            for (int e = 0; e < E; e++) vv[c][e] = z[RW::sigma(e)];
        }
    };
    C2<T> (&vu)[3][E] = *reinterpret_cast<C2<T>(*)[3][E]>(&v[0][0]);
    C2<T> (&vw)[3][E] = *reinterpret_cast<C2<T>(*)[3][E]>(&v[3][0]);
    load(0);
    unpack(vu);
    load(3);
    rows_w_fft<T, N, +1, 3>(vu, reg, tw, g);
    unpack(vw);
    rows_w_fft<T, N, +1, 3>(vw, reg, tw, g);
    if (ru) {
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int x = g + G * RW::tau(e);
                ru[c * 2 * PS + rrow<N>(2 * q) + x] = v[c][e].x;
                ru[c * 2 * PS + rrow<N>(2 * q) + N + x] = v[c][e].y;
            }
    }
    // (u x omega) / N^3 on rows g + G tau(e), reordered to the forward FFT's
    // input rows g + G sigma(e)
    C2<T> f[3][E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int eo = RW::tau_inv(RW::sigma(e));
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            const C2<T> u1 = v[c1][eo], u2 = v[c2][eo], w1 = v[3 + c1][eo], w2 = v[3 + c2][eo];
            const T re = u1.x * w2.x - u2.x * w1.x;
            const T im = u1.y * w2.y - u2.y * w1.y;
            f[c][e] = mk<T>(re * inp, im * inp);
        }
    }
    rows_w_fft<T, N, -1, 3>(f, reg, tw, g);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        // pack (as rows_pack_forward): A[k] = (Z[k] + conj Z[N-k]) / 2,
        // B[k] = (Z[k] - conj Z[N-k]) / (2 i), k = g + G m, m < E / 2
#pragma unroll
        for (int m = 0; m < H; m++) {
            const C2<T> Z = f[c][RW::tau_inv(m)];
            C2<T> M = shfl2<T>(f[c][RW::tau_inv(E - 1 - m)], partner);
            if (g == 0) M = f[c][RW::tau_inv((E - m) % E)];
            data[c * PS + ra + g + G * m] = mk<T>(T(0.5) * (Z.x + M.x), T(0.5) * (Z.y - M.y));
            data[c * PS + rb + g + G * m] = mk<T>(T(0.5) * (Z.y + M.y), T(-0.5) * (Z.x - M.x));
        }
    }
}

// c2r x pass of NC planes into real planes out ([y][z][x] each), components
// in groups of up to 3.  out may equal data (in place: a real row occupies
// the bytes of the spectral row with the same index, and a block reads all
// of its rows before writing them).
template <typename T, int N, int NC>
__global__ void __launch_bounds__((Cfg<T, N>::NT_ROW))
k_rows_c2r(const C2<T>* data, T* out, const C2<T>* __restrict__ tw) {
    using CF = Cfg<T, N>;
    constexpr int NG = NC < 3 ? NC : 3;
    constexpr int RP = CF::RP, NT = CF::NT_ROW, NXH = CF::NXH, LS = NG * RP + 1;
    constexpr size_t PS = plane_size(N), YS = row_stride(N);
    static_assert(NC % NG == 0, "component groups");
    __shared__ C2<T> lds[N * LS];
    const int tid = threadIdx.x;
    const int q0 = blockIdx.x * RP;
    for (int c0 = 0; c0 < NC; c0 += NG) {
        {
            const C2<T>* ra[NG * RP];
            const C2<T>* rb[NG * RP];
#pragma unroll
            for (int c = 0; c < NG; c++)
#pragma unroll
                for (int p = 0; p < RP; p++) {
                    const size_t row = size_t(q0 + p) * 2;
                    ra[c * RP + p] = data + (c0 + c) * PS + crow<N>(row);
                    rb[c * RP + p] = data + (c0 + c) * PS + crow<N>(row) + NXH;
                }
            rows_unpack_inverse<T, N, LS, NG * RP, NT>(lds, ra, rb, tid);
        }
        __syncthreads();
        fft_tile<T, N, NG * RP, LS, NT, CF::RMAX, +1>(lds, tw, tid);
        for (int e = tid; e < N * NG * RP; e += NT) {  // consecutive lanes: consecutive x
            const int col = e / N, nn = e % N;
            const int c = col / RP, p = col % RP;
            const size_t row = size_t(q0 + p) * 2;
            const C2<T> v = lds[nn * LS + col];
            out[(c0 + c) * 2 * PS + rrow<N>(row) + nn] = v.x;
            out[(c0 + c) * 2 * PS + rrow<N>(row) + N + nn] = v.y;
        }
        __syncthreads();
    }
}

// r2c x pass of NC real planes ([y][z][x]) into spectral rows (kx < N/2),
// components in groups of up to 3; out may equal data (in place).
template <typename T, int N, int NC>
__global__ void __launch_bounds__((Cfg<T, N>::NT_ROW))
k_rows_r2c(const T* data, C2<T>* out, const C2<T>* __restrict__ tw) {
    using CF = Cfg<T, N>;
    constexpr int NG = NC < 3 ? NC : 3;
    constexpr int RP = CF::RP, NT = CF::NT_ROW, NXH = CF::NXH, LS = NG * RP + 1;
    constexpr size_t PS = plane_size(N), YS = row_stride(N);
    static_assert(NC % NG == 0, "component groups");
    __shared__ C2<T> lds[N * LS];
    const int tid = threadIdx.x;
    const int q0 = blockIdx.x * RP;
    for (int c0 = 0; c0 < NC; c0 += NG) {
        for (int e = tid; e < N * NG * RP; e += NT) {  // consecutive lanes: consecutive x
            const int col = e / N, nn = e % N;
            const int c = col / RP, p = col % RP;
            const size_t row = size_t(q0 + p) * 2;
            lds[nn * LS + col] = mk<T>(data[(c0 + c) * 2 * PS + rrow<N>(row) + nn],
                                       data[(c0 + c) * 2 * PS + rrow<N>(row) + N + nn]);
        }
        __syncthreads();
        fft_tile<T, N, NG * RP, LS, NT, CF::RMAX, -1>(lds, tw, tid);
        for (int c = 0; c < NG; c++)
            for (int p = 0; p < RP; p++) {
                const size_t row = size_t(q0 + p) * 2;
                rows_pack_forward<T, N, LS>(lds, c * RP + p, out + (c0 + c) * PS + crow<N>(row),
                                            out + (c0 + c) * PS + crow<N>(row) + NXH, tid, NT);
            }
        __syncthreads();
    }
}

// ------------------------------------------------------------ last forward pass + update
// STAGE 1..3 of the integrating-factor SSP-RK3 step
// (vorticity_equation::step, bfps/cpp/vorticity_equation.cpp:358-439), fused
// with the tail of omega_nonlin (:323-355): dealias (kspace.cpp:461-485),
// curl, forcing (:241-288), force_divfree (kspace.cpp:487-519).
// data: planes 0..2 hold the output of the z forward pass ([y][kz][kx]).
// v0 = omega(t); vp = previous stage (stage 2: v1, stage 3: v2); dst receives
// the new stage (stage 3 writes v0 in place and applies the final projection).
template <typename T, int N, int STAGE, bool BATCH, int E_ = Cfg<T, N>::E_Y3, int B_ = Cfg<T, N>::B_Y3>
__global__ void __launch_bounds__(((N / E_) * B_))
k_fwd_y_update(const C2<T>* __restrict__ data, const C2<T>* v0, const C2<T>* __restrict__ vp,
               C2<T>* dst, const C2<T>* __restrict__ tw, Geom g, StepParams sp0, size_t bstride,
               const StepParams* __restrict__ spb) {
    using CF = Cfg<T, N>;
    // BATCH: box blockIdx.z of a batch, its fields bstride elements apart and
    // its own physics parameters in spb (a separate instantiation keeps the
    // single-box kernel's registers free of the batch bookkeeping)
    if constexpr (BATCH) {
        data += blockIdx.z * bstride;
        v0 += blockIdx.z * bstride;
        vp += blockIdx.z * bstride;
        dst += blockIdx.z * bstride;
    }
    const StepParams sp = BATCH ? spb[blockIdx.z] : sp0;
    constexpr int B = B_, E = E_, NT = (N / E_) * B_, NXH = CF::NXH;
    constexpr size_t PS = plane_size(N), YS = row_stride(N);
    using F = RegFft<T, N, B, B, NT, -1>;
    __shared__ C2<T> lds[N * B];
    const int tid = threadIdx.x;
    const int b = tid % B, grp = tid / B;
    // tiles narrower than a 128-byte line (1024^3): neighbouring tiles, the
    // two halves of each line, run on the same XCD and share its L2
    int tile = blockIdx.x, kz = blockIdx.y;
    if constexpr (B * sizeof(C2<T>) < 128) xcd_tile(tile, kz);
    const int kxi = tile * B + b;
    C2<T> res[3][E];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const C2<T>* src = data + c * PS + size_t(kz) * NXH + kxi;
#pragma unroll
        for (int e = 0; e < E; e++) res[c][e] = src[size_t(F::in_row(grp, e)) * YS];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) F::run(res[c], lds, tw, grp, b);
    const double kx = kxi * g.dkx, kzv = kwave(kz, N, g.dkz);
    const double ikM2 = 1.0 / g.kM2;
#pragma unroll
    for (int i = 0; i < E; i++) {
        const int ky = F::out_row(grp, i);
        const double kyv = kwave(ky, N, g.dky);
        const double k2 = kx * kx + kyv * kyv + kzv * kzv;
        const size_t off = size_t(ky) * YS + size_t(kz) * NXH + kxi;
        // dealias
        C2<T> a[3];
        if (g.smooth) {
            const double r = k2 * ikM2;
            const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4, r16 = r8 * r8;
            const double flt = exp(-36.0 * (r16 * r2));
#pragma unroll
            for (int c = 0; c < 3; c++) a[c] = mk<T>(T(res[c][i].x * flt), T(res[c][i].y * flt));
        } else {
#pragma unroll
            for (int c = 0; c < 3; c++) a[c] = k2 >= g.kM2 ? mk<T>(T(0), T(0)) : res[c][i];
        }
        // curl: i k x a
        C2<T> cu[3];
        cu[0] = mk<T>(T(-(kyv * a[2].y - kzv * a[1].y)), T(kyv * a[2].x - kzv * a[1].x));
        cu[1] = mk<T>(T(-(kzv * a[0].y - kx * a[2].y)), T(kzv * a[0].x - kx * a[2].x));
        cu[2] = mk<T>(T(-(kx * a[1].y - kyv * a[0].y)), T(kx * a[1].x - kyv * a[0].x));
        // forcing with the stage's source vorticity
        const C2<T>* srcv = STAGE == 1 ? v0 : vp;
        if (sp.forcing == 1 || sp.forcing == 3) {
            const double kn = sqrt(k2);
            if (sp.fk0 <= kn && sp.fk1 >= kn) {
                const double amp = sp.forcing == 1 ? sp.famp : sp.famp_dev[STAGE - 1];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const C2<T> s = srcv[c * PS + off];
                    cu[c] = mk<T>(T(cu[c].x + amp * double(s.x) * 1.0), T(cu[c].y + amp * double(s.y) * 1.0));
                }
            }
        } else if (sp.forcing == 2) {
            if (kxi == 0 && kz == 0 && (ky == sp.fmode || ky == N - sp.fmode))
                cu[2].x = T(cu[2].x - sp.famp * 1.0 / 2);
        }
        // force_divfree
        if (k2 > 0) {
            const double ik2 = 1.0 / k2;  // one division, then products
            const T t0 = T((kx * cu[0].x + kyv * cu[1].x + kzv * cu[2].x) * ik2);
            const T t1 = T((kx * cu[0].y + kyv * cu[1].y + kzv * cu[2].y) * ik2);
            cu[0] = mk<T>(T(cu[0].x - t0 * kx), T(cu[0].y - t1 * kx));
            cu[1] = mk<T>(T(cu[1].x - t0 * kyv), T(cu[1].y - t1 * kyv));
            cu[2] = mk<T>(T(cu[2].x - t0 * kzv), T(cu[2].y - t1 * kzv));
        } else {
            cu[0] = cu[1] = cu[2] = mk<T>(T(0), T(0));
        }
        // integrating-factor RK stage
        const double dt = sp.dt;
        if (STAGE == 1) {
            if (k2 <= g.kM2) {
                const double f0 = exp(-sp.nu * k2 * dt);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const C2<T> w0 = v0[c * PS + off];
                    dst[c * PS + off] = mk<T>(T((w0.x + dt * cu[c].x) * f0), T((w0.y + dt * cu[c].y) * f0));
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) dst[c * PS + off] = mk<T>(T(0), T(0));
            }
        } else if (STAGE == 2) {
            if (k2 <= g.kM2) {
                const double f0 = exp(-sp.nu * k2 * dt / 2);
                const double f1 = exp(sp.nu * k2 * dt / 2);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const C2<T> w0 = v0[c * PS + off];
                    const C2<T> w1 = vp[c * PS + off];
                    dst[c * PS + off] = mk<T>(T((3 * w0.x * f0 + (w1.x + dt * cu[c].x) * f1) * 0.25),
                                              T((3 * w0.y * f0 + (w1.y + dt * cu[c].y) * f1) * 0.25));
                }
            }
        } else {
            C2<T> nv[3];
            if (k2 <= g.kM2) {
                const double f0 = exp(-sp.nu * k2 * dt * 0.5);
                const double f3 = f0 * (1.0 / 3);  // products instead of six divisions by 3
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const C2<T> w0 = v0[c * PS + off];
                    const C2<T> w2 = vp[c * PS + off];
                    nv[c] = mk<T>(T((w0.x * f0 + 2 * (w2.x + dt * cu[c].x)) * f3),
                                  T((w0.y * f0 + 2 * (w2.y + dt * cu[c].y)) * f3));
                }
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++) nv[c] = v0[c * PS + off];
            }
            // final force_divfree of the new vorticity
            if (k2 > 0) {
                const double ik2 = 1.0 / k2;
                const T t0 = T((kx * nv[0].x + kyv * nv[1].x + kzv * nv[2].x) * ik2);
                const T t1 = T((kx * nv[0].y + kyv * nv[1].y + kzv * nv[2].y) * ik2);
                nv[0] = mk<T>(T(nv[0].x - t0 * kx), T(nv[0].y - t1 * kx));
                nv[1] = mk<T>(T(nv[1].x - t0 * kyv), T(nv[1].y - t1 * kyv));
                nv[2] = mk<T>(T(nv[2].x - t0 * kzv), T(nv[2].y - t1 * kzv));
            } else {
                nv[0] = nv[1] = nv[2] = mk<T>(T(0), T(0));
            }
#pragma unroll
            for (int c = 0; c < 3; c++) dst[c * PS + off] = nv[c];
        }
    }
}

// Fixed energy-injection-rate forcing (not a reference forcing type): the
// linear band forcing with amplitude eps / (2 E_band), E_band = 1/2 sum over
// fk0 <= |k| <= fk1 of nxmodes |omega_hat|^2 / k^2 (= 1/2 sum nxmodes |u_hat|^2
// for solenoidal omega), so that the band gains energy at the rate eps.  One
// block scans the box |k_i| <= fk1 and writes the amplitude to *amp.
template <typename T, int N>
__global__ void __launch_bounds__(256) k_band_amplitude(const C2<T>* __restrict__ w, double* amp, Geom g, double eps,
                                                        double fk0, double fk1, size_t bstride,
                                                        const StepParams* __restrict__ spb, int stage) {
    // box blockIdx.x of a batch: its source field, its own rate, band and
    // amplitude slot (spb[box].famp_dev + stage - 1)
    if (spb) {
        const StepParams& p = spb[blockIdx.x];
        if (p.forcing != 3) return;
        w += blockIdx.x * bstride;
        amp = const_cast<double*>(p.famp_dev) + (stage - 1);
        eps = p.famp;
        fk0 = p.fk0;
        fk1 = p.fk1;
    }
    constexpr int NXH = N / 2;
    constexpr size_t PS = plane_size(N), YS = row_stride(N);
    __shared__ double part[256];
    const int KX = min(NXH - 1, int(fk1 / g.dkx));
    const int KY = min(N / 2 - 1, int(fk1 / g.dky)), KZ = min(N / 2 - 1, int(fk1 / g.dkz));
    const int nyv = 2 * KY + 1, nzv = 2 * KZ + 1;
    const int total = (KX + 1) * nyv * nzv;
    double e = 0.0;
    for (int t = threadIdx.x; t < total; t += 256) {
        const int x = t % (KX + 1), zz = (t / (KX + 1)) % nzv - KZ, yy = t / ((KX + 1) * nzv) - KY;
        const double kx = x * g.dkx, ky = yy * g.dky, kz = zz * g.dkz;
        const double k2 = kx * kx + ky * ky + kz * kz;
        const double kn = sqrt(k2);
        if (!(fk0 <= kn && fk1 >= kn && k2 > 0)) continue;
        const size_t off = size_t(yy < 0 ? yy + N : yy) * YS + size_t(zz < 0 ? zz + N : zz) * NXH + x;
        double m = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const C2<T> v = w[c * PS + off];
            m += double(v.x) * double(v.x) + double(v.y) * double(v.y);
        }
        e += (x == 0 ? 1.0 : 2.0) * m / k2;
    }
    part[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double eb = 0.5 * part[0];
        *amp = eb > 0 ? eps / (2 * eb) : 0.0;
    }
}

// ------------------------------------------------------------ small spectral kernels
// field::symmetrize on the kx = 0 plane (bfps/cpp/field.cpp:900-978):
// Im(0,0,0) = 0; (0, kz) = conj(0, N-kz) for N/2 < kz < N;
// (ky, kz) = conj(N-ky, (N-kz) mod N) for N/2 < ky < N.
template <typename T>
__global__ void k_symmetrize(C2<T>* v, int n, int nc, size_t bstride) {
    v += blockIdx.y * bstride;  // box blockIdx.y of a batch
    const int nxh = n / 2;
    const size_t ps = plane_size(n), ys = row_stride(n);
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    const int ky = idx / n, kz = idx % n;
    for (int c = 0; c < nc; c++) {
        C2<T>* p = v + c * ps;
        if (ky == 0 && kz == 0) {
            p[0].y = T(0);
        } else if ((ky > n / 2) || (ky == 0 && kz > n / 2)) {
            const C2<T> m = p[size_t((n - ky) % n) * ys + size_t((n - kz) % n) * nxh];
            p[size_t(ky) * ys + size_t(kz) * nxh] = cconj<T>(m);
        }
    }
}

// low_pass(kM) (k^2 >= kM^2 -> 0) followed by force_divfree
// (io_checkpoint read projections, bfps/cpp/vorticity_equation.hpp:118-127)
template <typename T>
__global__ void k_lowpass_divfree(C2<T>* v, Geom g) {
    const int n = g.n, nxh = n / 2;
    const size_t ps = plane_size(n);
    const size_t lidx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lidx >= size_t(n) * n * nxh) return;
    const size_t idx = sphys(n, lidx);
    const int kxi = lidx % nxh, kz = (lidx / nxh) % n, ky = lidx / (size_t(nxh) * n);
    const double kx = kxi * g.dkx, kyv = kwave(ky, n, g.dky), kzv = kwave(kz, n, g.dkz);
    const double k2 = kx * kx + kyv * kyv + kzv * kzv;
    C2<T> a[3];
    for (int c = 0; c < 3; c++) a[c] = v[c * ps + idx];
    if (k2 >= g.kM2) a[0] = a[1] = a[2] = mk<T>(T(0), T(0));
    if (k2 > 0) {
        const double ik2 = 1.0 / k2;  // as the step's projections
        const T t0 = T((kx * a[0].x + kyv * a[1].x + kzv * a[2].x) * ik2);
        const T t1 = T((kx * a[0].y + kyv * a[1].y + kzv * a[2].y) * ik2);
        a[0] = mk<T>(T(a[0].x - t0 * kx), T(a[0].y - t1 * kx));
        a[1] = mk<T>(T(a[1].x - t0 * kyv), T(a[1].y - t1 * kyv));
        a[2] = mk<T>(T(a[2].x - t0 * kzv), T(a[2].y - t1 * kzv));
    } else {
        a[0] = a[1] = a[2] = mk<T>(T(0), T(0));
    }
    for (int c = 0; c < 3; c++) v[c * ps + idx] = a[c];
}

// Biot-Savart into a spectral field (compute_velocity without symmetrize;
// the caller symmetrizes), used for host-visible cvelocity.
template <typename T>
__global__ void k_biot_savart(const C2<T>* __restrict__ w, C2<T>* __restrict__ u, Geom g) {
    const int n = g.n, nxh = n / 2;
    const size_t ps = plane_size(n);
    const size_t lidx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lidx >= size_t(n) * n * nxh) return;
    const size_t idx = sphys(n, lidx);
    const int kxi = lidx % nxh, kz = (lidx / nxh) % n, ky = lidx / (size_t(nxh) * n);
    const double kx = kxi * g.dkx, kyv = kwave(ky, n, g.dky), kzv = kwave(kz, n, g.dkz);
    const double k2 = kx * kx + kyv * kyv + kzv * kzv;
    const C2<T> a0 = w[idx], a1 = w[ps + idx], a2 = w[2 * ps + idx];
    if (k2 <= g.kM2 && k2 > 0) {
        const double ik2 = 1.0 / k2;  // as the step's fused Biot-Savart pass
        u[idx] = mk<T>(T(-(kyv * a2.y - kzv * a1.y) * ik2), T((kyv * a2.x - kzv * a1.x) * ik2));
        u[ps + idx] = mk<T>(T(-(kzv * a0.y - kx * a2.y) * ik2), T((kzv * a0.x - kx * a2.x) * ik2));
        u[2 * ps + idx] = mk<T>(T(-(kx * a1.y - kyv * a0.y) * ik2), T((kx * a1.x - kyv * a0.x) * ik2));
    } else {
        u[idx] = u[ps + idx] = u[2 * ps + idx] = mk<T>(T(0), T(0));
    }
}

// compute_vorticity (bfps/cpp/vorticity_equation.cpp:171-204): the curl
// w = i k x u for k^2 <= kM^2 (the k = 0 mode included), zero beyond; the
// caller symmetrizes.
template <typename T>
__global__ void k_curl(const C2<T>* __restrict__ u, C2<T>* __restrict__ w, Geom g) {
    const int n = g.n, nxh = n / 2;
    const size_t ps = plane_size(n);
    const size_t lidx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lidx >= size_t(n) * n * nxh) return;
    const size_t idx = sphys(n, lidx);
    const int kxi = lidx % nxh, kz = (lidx / nxh) % n, ky = lidx / (size_t(nxh) * n);
    const double kx = kxi * g.dkx, kyv = kwave(ky, n, g.dky), kzv = kwave(kz, n, g.dkz);
    const double k2 = kx * kx + kyv * kyv + kzv * kzv;
    const C2<T> a0 = u[idx], a1 = u[ps + idx], a2 = u[2 * ps + idx];
    if (k2 <= g.kM2) {
        w[idx] = mk<T>(T(-(kyv * a2.y - kzv * a1.y)), T(kyv * a2.x - kzv * a1.x));
        w[ps + idx] = mk<T>(T(-(kzv * a0.y - kx * a2.y)), T(kzv * a0.x - kx * a2.x));
        w[2 * ps + idx] = mk<T>(T(-(kx * a1.y - kyv * a0.y)), T(kx * a1.x - kyv * a0.x));
    } else {
        w[idx] = w[ps + idx] = w[2 * ps + idx] = mk<T>(T(0), T(0));
    }
}

// compute_gradient (bfps/cpp/field.cpp:1042-1101) of an ncs-component
// spectral field (ncs = 1 or 3): dst component j ncs + i = i k_j src_i for
// k^2 < kM^2 (strict), zero elsewhere.
template <typename T>
__global__ void k_gradient(const C2<T>* __restrict__ src, C2<T>* __restrict__ dst, Geom g, int ncs) {
    const int n = g.n, nxh = n / 2;
    const size_t ps = plane_size(n);
    const size_t lidx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lidx >= size_t(n) * n * nxh) return;
    const size_t idx = sphys(n, lidx);
    const int kxi = lidx % nxh, kz = (lidx / nxh) % n, ky = lidx / (size_t(nxh) * n);
    const double kv[3] = {kxi * g.dkx, kwave(ky, n, g.dky), kwave(kz, n, g.dkz)};
    const bool in = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2] < g.kM2;
    for (int i = 0; i < ncs; i++) {
        const C2<T> a = src[i * ps + idx];
        for (int j = 0; j < 3; j++)
            dst[(j * ncs + i) * ps + idx] = in ? mk<T>(T(-kv[j] * a.y), T(kv[j] * a.x)) : mk<T>(T(0), T(0));
    }
}

// kspace::cospectrum (bfps/cpp/kspace.cpp:521-598): FP64 shell sums
// S[s][i][j] += nxmodes (Re a_i Re b_j + Im a_i Im b_j), s = int(|k| / dk),
// k^2 <= kM^2, nxmodes = 1 at kx = 0 else 2.
template <typename T>
__global__ void k_cospectrum(const C2<T>* __restrict__ a, const C2<T>* __restrict__ bb, double* out,
                             Geom g, double dk, int nshells) {
    extern __shared__ double acc[];
    const int n = g.n, nxh = n / 2;
    const size_t ps = plane_size(n), nlog = size_t(n) * n * nxh;
    for (int i = threadIdx.x; i < nshells * 9; i += blockDim.x) acc[i] = 0.0;
    __syncthreads();
    for (size_t lidx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; lidx < nlog; lidx += size_t(gridDim.x) * blockDim.x) {
        const int kxi = lidx % nxh, kz = (lidx / nxh) % n, ky = lidx / (size_t(nxh) * n);
        const size_t idx = sphys(n, lidx);
        const double kx = kxi * g.dkx, kyv = kwave(ky, n, g.dky), kzv = kwave(kz, n, g.dkz);
        const double k2 = kx * kx + kyv * kyv + kzv * kzv;
        if (k2 > g.kM2) continue;
        const double nxm = kxi == 0 ? 1.0 : 2.0;
        const int s = int(sqrt(k2) / dk) * 9;
        C2<T> av[3], bv[3];
        for (int c = 0; c < 3; c++) { av[c] = a[c * ps + idx]; bv[c] = bb[c * ps + idx]; }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                atomicAdd(&acc[s + i * 3 + j], nxm * (double(av[i].x) * double(bv[j].x) + double(av[i].y) * double(bv[j].y)));
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nshells * 9; i += blockDim.x)
        if (acc[i] != 0.0) atomicAdd(&out[i], acc[i]);
}

}  // namespace bfpsh
