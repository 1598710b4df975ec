// Statistics either side of the time step (NSVE::do_stats and
// NSVEparticles::do_stats, bfps/cpp/full_code/NSVE.cpp:101-135,
// NSVEparticles.cpp:70-98): pressure and Lagrangian acceleration
// (bfps/cpp/vorticity_equation.cpp:441-578) and the real-space moments and
// histograms of field::compute_rspace_stats (bfps/cpp/field.cpp:661-850).
// All of them are single HBM sweeps (or, for the pressure, two vector
// transforms around them).
#pragma once
#include "kernels.hpp"

namespace bfpsh {

// Products of the real velocity (planar [c][y][z][x], n3 reals per plane
// including the row pads, whose products are never read)
// for the pressure: pass 0 the diagonal u_c u_c, pass 1 u_c u_{c+1}
// (vorticity_equation.cpp:451-459,484-492), in the field precision.
template <typename T>
__global__ void k_uu(const T* __restrict__ u, T* __restrict__ out, size_t n3, int pass) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n3; i += size_t(gridDim.x) * blockDim.x) {
        const T a0 = u[i], a1 = u[n3 + i], a2 = u[2 * n3 + i];
        if (pass == 0) {
            out[i] = a0 * a0;
            out[n3 + i] = a1 * a1;
            out[2 * n3 + i] = a2 * a2;
        } else {
            out[i] = a0 * a1;
            out[n3 + i] = a1 * a2;
            out[2 * n3 + i] = a2 * a0;
        }
    }
}

// One pass of the pressure from the transformed products f (3 spectral
// planes): dealias f (kspace::dealias, kspace.cpp:461-485), then
// pass 0: p = -(kx^2 f0 + ky^2 f1 + kz^2 f2) for 0 < k^2 <= kM^2, else 0;
// pass 1: p -= 2 (kx ky f0 + ky kz f1 + kz kx f2), p /= npoints k^2
// (vorticity_equation.cpp:462-482,495-518).
template <typename T>
__global__ void k_pressure(const C2<T>* __restrict__ f, C2<T>* __restrict__ p, Geom g, int pass) {
    const int n = g.n, nxh = n / 2;
    const size_t ps = plane_size(n);
    const size_t lidx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lidx >= size_t(n) * n * nxh) return;
    const size_t idx = sphys(n, lidx);
    const int kxi = lidx % nxh, kz = (lidx / nxh) % n, ky = lidx / (size_t(nxh) * n);
    const double kx = kxi * g.dkx, kyv = kwave(ky, n, g.dky), kzv = kwave(kz, n, g.dkz);
    const double k2 = kx * kx + kyv * kyv + kzv * kzv;
    if (!(k2 <= g.kM2 && k2 > 0)) {
        if (pass == 0) p[idx] = mk<T>(T(0), T(0));
        return;
    }
    C2<T> a[3];
#pragma unroll
    for (int c = 0; c < 3; c++) a[c] = f[c * ps + idx];
    if (g.smooth) {
        const double r = k2 / g.kM2;
        const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4, r16 = r8 * r8;
        const double flt = exp(-36.0 * (r16 * r2));
#pragma unroll
        for (int c = 0; c < 3; c++) a[c] = mk<T>(T(a[c].x * flt), T(a[c].y * flt));
    } else if (k2 >= g.kM2) {
        a[0] = a[1] = a[2] = mk<T>(T(0), T(0));
    }
    if (pass == 0) {
        p[idx] = mk<T>(T(-(kx * kx * a[0].x + kyv * kyv * a[1].x + kzv * kzv * a[2].x)),
                       T(-(kx * kx * a[0].y + kyv * kyv * a[1].y + kzv * kzv * a[2].y)));
    } else {
        const double npoints = double(n) * n * n;
        C2<T> q = p[idx];
        q.x = T(q.x - 2 * (kx * kyv * a[0].x + kyv * kzv * a[1].x + kzv * kx * a[2].x));
        q.y = T(q.y - 2 * (kx * kyv * a[0].y + kyv * kzv * a[1].y + kzv * kx * a[2].y));
        q.x = T(q.x / (npoints * k2));
        q.y = T(q.y / (npoints * k2));
        p[idx] = q;
    }
}

// Lagrangian acceleration from the spectral velocity u (3 planes, Biot-Savart
// of cvorticity, symmetrized) and the pressure p: for k^2 <= kM^2
// a = -nu k^2 u (+ famplitude u in fk0 <= |k| <= fk1 for linear forcing)
// - i k p, else 0 (vorticity_equation.cpp:540-576).
template <typename T>
__global__ void k_lagr_acc(const C2<T>* __restrict__ u, const C2<T>* __restrict__ p, C2<T>* __restrict__ acc,
                           Geom g, double nu, int linear, double famp, double fk0, double fk1) {
    const int n = g.n, nxh = n / 2;
    const size_t ps = plane_size(n);
    const size_t lidx = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lidx >= size_t(n) * n * nxh) return;
    const size_t idx = sphys(n, lidx);
    const int kxi = lidx % nxh, kz = (lidx / nxh) % n, ky = lidx / (size_t(nxh) * n);
    const double kv[3] = {kxi * g.dkx, kwave(ky, n, g.dky), kwave(kz, n, g.dkz)};
    const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
    C2<T> a[3] = {mk<T>(T(0), T(0)), mk<T>(T(0), T(0)), mk<T>(T(0), T(0))};
    if (k2 <= g.kM2) {
        C2<T> uu[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            uu[c] = u[c * ps + idx];
            a[c] = mk<T>(T(-nu * k2 * uu[c].x), T(-nu * k2 * uu[c].y));
        }
        if (linear) {
            const double kn = sqrt(k2);
            if (fk0 <= kn && fk1 >= kn) {
#pragma unroll
                for (int c = 0; c < 3; c++) a[c] = mk<T>(T(a[c].x + famp * uu[c].x), T(a[c].y + famp * uu[c].y));
            }
        }
        const C2<T> q = p[idx];
#pragma unroll
        for (int c = 0; c < 3; c++) a[c] = mk<T>(T(a[c].x + kv[c] * q.y), T(a[c].y - kv[c] * q.x));
    }
#pragma unroll
    for (int c = 0; c < 3; c++) acc[c * ps + idx] = a[c];
}

// field::normalize (field.cpp:891-898): data /= npoints in the field
// precision
template <typename T>
__global__ void k_div(T* __restrict__ d, size_t count, T npoints) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x)
        d[i] = d[i] / npoints;
}

// Moments and histograms of a real field (field::compute_rspace_stats).
// NV = 4: the three components of a vector field and its magnitude; NV = 1:
// component blockIdx.y of a 1- or 9-component field.  Per value: min and
// max (starting at 0; the magnitude's min at its estimate), sums of v^1..v^8
// in FP64, histogram bins of width 2 M / nbins from -M (the magnitude's from
// 0 with half that width), counted in LDS ([value][bin], so that the lanes
// of a wave spread over the banks) and added to hist[nbins][nvals].
// Each workgroup writes its partial moments to part[group][block][10][NV].
template <typename T, int NV>
__global__ void __launch_bounds__(256)
k_rstats(const T* __restrict__ d, int n, int nvals, int nbins, const double* __restrict__ maxest,
         double* __restrict__ part, unsigned long long* __restrict__ hist) {
    // every product and sum rounded on its own, as the reference's loop (no
    // fused multiply-adds)
#pragma clang fp contract(off)
    extern __shared__ unsigned int lh[];  // [NV][nbins]
    __shared__ double red[4][10 * NV];
    constexpr int NCV = NV == 4 ? 3 : 1;  // components read per point
    const int c0 = NV == 4 ? 0 : blockIdx.y;
    const int tid = threadIdx.x;
    for (int i = tid; i < nbins * NV; i += blockDim.x) lh[i] = 0u;
    __syncthreads();
    double me[NV], bs[NV], mn[NV], mx[NV], s[8][NV];
#pragma unroll
    for (int v = 0; v < NV; v++) {
        me[v] = maxest[c0 + v];
        bs[v] = 2 * me[v] / nbins;
        mn[v] = 0.0;
        mx[v] = 0.0;
#pragma unroll
        for (int k = 0; k < 8; k++) s[k][v] = 0.0;
    }
    if (NV == 4) mn[3] = me[3];
    // the loads of the next PF points of this thread are in flight while a
    // point is processed (the per-point work is short against the memory
    // latency at the occupancy the FP64 accumulators allow)
    constexpr int PF = 4;
    // logical point j = y n^2 + (z n + x) of a plane lies at reals
    // y ys2 + (z n + x); planes cps reals apart.  The load cursor keeps
    // (y, z n + x) of the next point to load and advances by the grid
    // stride without a division per point.
    const size_t n3 = size_t(n) * n * n, cps = 2 * plane_size(n), ys2 = 2 * row_stride(n);
    const uint32_t n2 = uint32_t(n) * uint32_t(n);
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    const size_t i0 = size_t(blockIdx.x) * blockDim.x + tid;
    const uint32_t sq = uint32_t(stride / n2), sr = uint32_t(stride % n2);
    uint32_t lq = uint32_t(i0 / n2), lr = uint32_t(i0 % n2);
    size_t lj = i0;  // logical index at the cursor
    auto load_next = [&](T (&dst)[NCV]) {
        const size_t off = size_t(lq) * ys2 + lr;
#pragma unroll
        for (int c = 0; c < NCV; c++) dst[c] = lj < n3 ? d[size_t(c0 + c) * cps + off] : T(0);
        lj += stride;
        lq += sq;
        lr += sr;
        if (lr >= n2) {
            lr -= n2;
            lq++;
        }
    };
    T buf[PF][NCV];
#pragma unroll
    for (int p = 0; p < PF; p++) load_next(buf[p]);
    for (size_t i = i0; i < n3; i += stride) {
        double val[NV];
#pragma unroll
        for (int c = 0; c < NCV; c++) val[c] = double(buf[0][c]);
#pragma unroll
        for (int p = 0; p + 1 < PF; p++)
#pragma unroll
            for (int c = 0; c < NCV; c++) buf[p][c] = buf[p + 1][c];
        load_next(buf[PF - 1]);  // the point PF strides ahead
        if (NV == 4) {
            // |u| = sqrt(((0 + u0^2) + u1^2) + u2^2)
            double m2 = val[0] * val[0];
            m2 = m2 + val[1] * val[1];
            m2 = m2 + val[2] * val[2];
            val[NV - 1] = sqrt(m2);
        }
#pragma unroll
        for (int v = 0; v < NV; v++) {
            mn[v] = val[v] < mn[v] ? val[v] : mn[v];
            mx[v] = val[v] > mx[v] ? val[v] : mx[v];
            const bool magnitude = NV == 4 && v == 3;
            const double fb = floor((magnitude ? val[v] * 2.0 : val[v] + me[v]) / bs[v]);
            if (fb >= 0.0 && fb < double(nbins)) atomicAdd(&lh[v * nbins + int(fb)], 1u);
            double pw = 1.0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                pw = val[v] * pw;
                s[k][v] = s[k][v] + pw;
            }
        }
    }
    // workgroup reduction: within each wave, then across the four waves
#pragma unroll
    for (int o = 32; o > 0; o /= 2) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
            mn[v] = fmin(mn[v], __shfl_xor(mn[v], o, 64));
            mx[v] = fmax(mx[v], __shfl_xor(mx[v], o, 64));
#pragma unroll
            for (int k = 0; k < 8; k++) s[k][v] += __shfl_xor(s[k][v], o, 64);
        }
    }
    const int wave = tid / 64, lane = tid % 64;
    if (lane == 0) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
            red[wave][v] = mn[v];
            red[wave][9 * NV + v] = mx[v];
#pragma unroll
            for (int k = 0; k < 8; k++) red[wave][(k + 1) * NV + v] = s[k][v];
        }
    }
    __syncthreads();
    if (tid < 10 * NV) {
        const int row = tid / NV;
        double r = red[0][tid];
        for (int w = 1; w < int(blockDim.x / 64); w++) {
            const double x = red[w][tid];
            r = row == 0 ? fmin(r, x) : (row == 9 ? fmax(r, x) : r + x);
        }
        part[(size_t(blockIdx.y) * gridDim.x + blockIdx.x) * 10 * NV + tid] = r;
    }
    for (int i = tid; i < nbins * NV; i += blockDim.x) {
        const unsigned int cnt = lh[i];
        if (cnt) atomicAdd(&hist[size_t(i % nbins) * nvals + c0 + i / nbins], (unsigned long long)cnt);
    }
}

// Combine the partial moments of nblk workgroups (in block order) into
// out[10][nvals]; rows 1..8 become means.  One workgroup per component group.
template <int NV>
__global__ void k_rstats_final(const double* __restrict__ part, int nblk, int nvals, double npoints,
                               double* __restrict__ out) {
    const int t = threadIdx.x;
    if (t >= 10 * NV) return;
    const int row = t / NV, v = t % NV;
    const int c0 = NV == 4 ? 0 : blockIdx.x;
    const double* pp = part + size_t(blockIdx.x) * nblk * 10 * NV;
    double r = pp[t];
    for (int b = 1; b < nblk; b++) {
        const double x = pp[size_t(b) * 10 * NV + t];
        r = row == 0 ? fmin(r, x) : (row == 9 ? fmax(r, x) : r + x);
    }
    if (row >= 1 && row <= 8) r /= npoints;
    out[row * nvals + c0 + v] = r;
}

}  // namespace bfpsh
