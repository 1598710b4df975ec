// bfps oracle -- TEST INFRASTRUCTURE ONLY.
//
// A single-process CPU restatement (C++17 + OpenMP) of the reference's
// vorticity-equation time step and tracer loop, used to check the HIP
// library and as the timed CPU baseline.  It is never linked into, loaded by,
// or called from the product library.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg use it.
//
// What it restates (reference file:line):
//   * field layouts: real  [z][y][x (nx+2 padded)][c]  (field.cpp:85-96)
//                    complex transposed [ky][kz][kx=nx/2+1][c] (field.cpp:97-101)
//   * unnormalised FFTW-MPI r2c (TRANSPOSED_OUT) / c2r (TRANSPOSED_IN)
//     (field.cpp:105-118,144-162), here with an own mixed-radix Stockham FFT
//   * kspace geometry, dealias filter, low_pass, force_divfree, cospectrum
//     (kspace.cpp:34-167,235-253,461-598; kspace.hpp:117-192)
//   * field::symmetrize (field.cpp:900-978), compute_gradient (field.cpp:1042-1101)
//   * vorticity_equation compute_vorticity/compute_velocity/add_forcing/
//     omega_nonlin/step (vorticity_equation.cpp:171-439), io_checkpoint read
//     projections (vorticity_equation.hpp:107-128)
//   * particles: get_norm_pos_in_cell / pbc_field_layer / apply_computation
//     (particles_field_computer.hpp:55-183), Adams-Bashforth 1..6
//     (particles_adams_bashforth.hpp:18-133), completeLoop order and rhs
//     rotation (particles_system.hpp:173-217), box geometry
//     (particles_system_builder.hpp:173-189)
//   * interpolation weights beta_Lagrange_n{n} / beta_n{n}_m{m}, rebuilt at
//     run time from their definition (Lagrange basis, two-point Hermite
//     matching of Lagrange derivatives) by numeric linear solves.
//   * statistics either side of the path: compute_pressure and
//     compute_Lagrangian_acceleration (vorticity_equation.cpp:441-578),
//     field::compute_rspace_stats moments and histograms (field.cpp:661-850)
//
// Arithmetic types follow the reference: field data in rnumber (float or
// double), wavenumbers / k^2 / factors in double, particle data in double.
//
// Parity status (each check lives under tests/):
//   * 3D FFT: numpy.fft (pocketfft) on seven shapes incl. radix 3 and 5,
//     brute-force DFT on 4^3 (test_oracle_fft.py)
//   * symmetrize: identity on Hermitian data, per-mode write pattern
//     (test_oracle_fft.py)
//   * whole NSVEparticles pipeline, SMOOTH, linear forcing, n1m1, AB4 with
//     its start-up ramp: the reference's golden checkpoint B32p1e4 (32^3,
//     64 steps, restart at 32), FP32 and FP64 (test_oracle_fixture.py)
//   * step / compute_velocity / gradient / forcing / projections /
//     cospectrum / kshell: closed forms -- Beltrami decay and analytic u and
//     grad u, RK3 stability polynomial of a forced single wave, Kolmogorov
//     forcing = Simpson's rule, low_pass + divfree, Parseval, lattice counts
//     (test_oracle_closed_forms.py)
//   * interpolation weights: the reference's spline_n*.cpp / Lagrange_polys.cpp
//     compiled from source, plus partition of unity and polynomial
//     reproduction (test_interp_weights.py); gather + AB1..AB6: convergence
//     to the analytic ABC field and order-k convergence of AB
//     (test_oracle_closed_forms.py)
//   * nonlinear term and dealiasing: the initial rate of a triad
//     k1 + k2 -> k3 against its Fourier-coefficient closed form, with k3 where
//     the SMOOTH filter is between 0.05 and 0.95, and across the TWO_THIRDS
//     cut (test_oracle_closed_forms.py)
//   * pressure and Lagrangian acceleration: the ABC flow, p = -|u|^2/2 +
//     const and a = grad |u|^2/2 - nu u, in FP64 and FP32; real-space
//     statistics: a constant field (min/max starting at 0), sin x (moments =
//     central binomial coefficients / 2^n, histogram of the eight values of
//     an 8-point period counted by hand) and a unit-magnitude field
//     (test_oracle_closed_forms.py)
//   * not reproducible bit for bit (last-bit effects, covered by
//     tolerances): the reference's FFTW plan-dependent and -ffast-math
//     rounding.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include <omp.h>

namespace orc {

// ------------------------------------------------------------------------ //
// 1D complex FFT: iterative Stockham autosort, mixed radix (4, 2, 3, 5, odd).
// Stage with radix R after sub-length L:
//   a_r = x[j + r*n/R] * w_{L R}^{r k},  k = j mod L
//   y[(j/L) L R + k + r L] = DFT_R(a)_r
// ------------------------------------------------------------------------ //
template <typename T>
struct Fft1d {
    int n;
    std::vector<int> radix;
    std::vector<std::complex<T>> tw_fwd;  // exp(-2 pi i t / n), t = 0..n-1
    explicit Fft1d(int n_) : n(n_) {
        int m = n;
        while (m % 4 == 0) { radix.push_back(4); m /= 4; }
        while (m % 2 == 0) { radix.push_back(2); m /= 2; }
        for (int p = 3; m > 1; p += 2)
            while (m % p == 0) { radix.push_back(p); m /= p; }
        tw_fwd.resize(n);
        for (int t = 0; t < n; t++) {
            long double a = -2.0L * 3.14159265358979323846264338327950288L * t / n;
            tw_fwd[t] = std::complex<T>(T(std::cos(a)), T(std::sin(a)));
        }
    }
    // in/out: x (length n), scratch: y (length n); sign = -1 forward, +1 backward
    void run(std::complex<T>* x, std::complex<T>* y, int sign) const {
        if (n == 1) return;
        std::complex<T>* src = x;
        std::complex<T>* dst = y;
        int L = 1;
        std::complex<T> a[64], b[64];
        std::vector<std::complex<T>> big;
        for (int R : radix) {
            const int nR = n / R;
            const int step = n / (L * R);
            for (int j = 0; j < nR; j++) {
                const int k = j % L;
                for (int r = 0; r < R; r++) {
                    std::complex<T> v = src[j + r * nR];
                    if (r && k) {
                        std::complex<T> w = tw_fwd[(size_t(r) * k * step) % n];
                        if (sign > 0) w = std::conj(w);
                        v *= w;
                    }
                    a[r] = v;
                }
                dft(a, b, R, sign);
                const int base = (j / L) * L * R + k;
                for (int r = 0; r < R; r++) dst[base + r * L] = b[r];
            }
            std::swap(src, dst);
            L *= R;
        }
        if (src != x) std::copy(src, src + n, x);
    }
    void dft(const std::complex<T>* a, std::complex<T>* b, int R, int sign) const {
        typedef std::complex<T> C;
        const T s = T(sign);
        switch (R) {
        case 2:
            b[0] = a[0] + a[1]; b[1] = a[0] - a[1];
            return;
        case 4: {
            C t0 = a[0] + a[2], t1 = a[0] - a[2];
            C t2 = a[1] + a[3], t3 = a[1] - a[3];
            C t3i(-s * t3.imag(), s * t3.real());  // i*sign*t3
            b[0] = t0 + t2; b[2] = t0 - t2;
            b[1] = t1 + t3i; b[3] = t1 - t3i;
            return;
        }
        default: {
            const int step = n / R;
            for (int q = 0; q < R; q++) {
                C acc = a[0];
                for (int r = 1; r < R; r++) {
                    C w = tw_fwd[(size_t(q) * r % R) * step];
                    if (sign > 0) w = std::conj(w);
                    acc += a[r] * w;
                }
                b[q] = acc;
            }
        }
        }
    }
};

// ------------------------------------------------------------------------ //
// 3D real FFT with the reference's layouts, nc interleaved components.
// ------------------------------------------------------------------------ //
template <typename T>
struct Fft3d {
    int nx, ny, nz, nc, nxh;
    Fft1d<T> fx, fy, fz;
    Fft3d(int nx_, int ny_, int nz_, int nc_)
        : nx(nx_), ny(ny_), nz(nz_), nc(nc_), nxh(nx_ / 2 + 1), fx(nx_), fy(ny_), fz(nz_) {}

    typedef std::complex<T> C;
    size_t rsize() const { return size_t(nz) * ny * (nx + 2) * nc; }

    // c2r of two Hermitian half rows A, B (length nxh, stride s) -> real rows
    void c2r_pair(const C* A, const C* B, size_t s, T* ra, T* rb, size_t rs, C* w, C* y) const {
        const int h = nx / 2;
        for (int k = 0; k <= h; k++) {
            C a = A[k * s], b = B ? B[k * s] : C(0, 0);
            if (k == 0 || k == h) { a = C(a.real(), 0); b = C(b.real(), 0); }
            w[k] = C(a.real() - b.imag(), a.imag() + b.real());
            if (k > 0 && k < h) {
                C ac = std::conj(a), bc = std::conj(b);
                w[nx - k] = C(ac.real() - bc.imag(), ac.imag() + bc.real());
            }
        }
        fx.run(w, y, +1);
        for (int i = 0; i < nx; i++) {
            ra[i * rs] = w[i].real();
            if (rb) rb[i * rs] = w[i].imag();
        }
    }
    void r2c_pair(const T* ra, const T* rb, size_t rs, C* A, C* B, size_t s, C* w, C* y) const {
        for (int i = 0; i < nx; i++) w[i] = C(ra[i * rs], rb ? rb[i * rs] : T(0));
        fx.run(w, y, -1);
        const int h = nx / 2;
        for (int k = 0; k <= h; k++) {
            C z = w[k], zc = std::conj(w[(nx - k) % nx]);
            A[k * s] = C((z.real() + zc.real()) * T(0.5), (z.imag() + zc.imag()) * T(0.5));
            if (B) {
                C d = z - zc;  // = 2 i B
                B[k * s] = C(d.imag() * T(0.5), -d.real() * T(0.5));
            }
        }
    }

    // in place on buf (size rsize()): complex [ky][kz][kx][c] -> real [z][y][x pad][c]
    void c2r(T* buf) const {
        C* cd = reinterpret_cast<C*>(buf);
        std::vector<C> tmp(size_t(nz) * ny * nxh * nc);
        const int nthreads = omp_get_max_threads();
        const int nmax = std::max(nx, std::max(ny, nz));
        std::vector<C> work(size_t(nthreads) * 2 * nmax);
        // along kz for each (ky, kx, c)
        #pragma omp parallel for schedule(static)
        for (long line = 0; line < long(ny) * nxh * nc; line++) {
            C* w = &work[size_t(omp_get_thread_num()) * 2 * nmax];
            const int c = line % nc, kx = (line / nc) % nxh, ky = line / (nc * nxh);
            C* base = cd + (size_t(ky) * nz * nxh + kx) * nc + c;
            const size_t st = size_t(nxh) * nc;
            for (int z = 0; z < nz; z++) w[z] = base[z * st];
            fz.run(w, w + nmax, +1);
            for (int z = 0; z < nz; z++) base[z * st] = w[z];
        }
        // transpose to [z][ky][kx][c], then along ky
        #pragma omp parallel for schedule(static)
        for (long zy = 0; zy < long(nz) * ny; zy++) {
            const int z = zy / ny, ky = zy % ny;
            std::copy(cd + (size_t(ky) * nz + z) * nxh * nc, cd + (size_t(ky) * nz + z + 1) * nxh * nc,
                      tmp.begin() + (size_t(z) * ny + ky) * nxh * nc);
        }
        #pragma omp parallel for schedule(static)
        for (long line = 0; line < long(nz) * nxh * nc; line++) {
            C* w = &work[size_t(omp_get_thread_num()) * 2 * nmax];
            const int c = line % nc, kx = (line / nc) % nxh, z = line / (nc * nxh);
            C* base = tmp.data() + (size_t(z) * ny * nxh + kx) * nc + c;
            const size_t st = size_t(nxh) * nc;
            for (int y = 0; y < ny; y++) w[y] = base[y * st];
            fy.run(w, w + nmax, +1);
            for (int y = 0; y < ny; y++) base[y * st] = w[y];
        }
        // c2r along x, rows paired
        const long nrows = long(nz) * ny * nc;
        #pragma omp parallel for schedule(static)
        for (long p = 0; p < (nrows + 1) / 2; p++) {
            C* w = &work[size_t(omp_get_thread_num()) * 2 * nmax];
            const long r0 = 2 * p, r1 = 2 * p + 1;
            auto rowc = [&](long r) { const int c = r % nc; const long zy = r / nc;
                                      return tmp.data() + size_t(zy) * nxh * nc + c; };
            auto rowr = [&](long r) { const int c = r % nc; const long zy = r / nc;
                                      return buf + size_t(zy) * (nx + 2) * nc + c; };
            c2r_pair(rowc(r0), r1 < nrows ? rowc(r1) : nullptr, nc, rowr(r0),
                     r1 < nrows ? rowr(r1) : nullptr, nc, w, w + nmax);
        }
        // padding entries: zero (unspecified in the reference)
        for (long zy = 0; zy < long(nz) * ny; zy++)
            for (int x = nx; x < nx + 2; x++)
                for (int c = 0; c < nc; c++) buf[(size_t(zy) * (nx + 2) + x) * nc + c] = 0;
    }

    // in place: real [z][y][x pad][c] -> complex [ky][kz][kx][c]
    void r2c(T* buf) const {
        C* cd = reinterpret_cast<C*>(buf);
        std::vector<C> tmp(size_t(nz) * ny * nxh * nc);
        const int nthreads = omp_get_max_threads();
        const int nmax = std::max(nx, std::max(ny, nz));
        std::vector<C> work(size_t(nthreads) * 2 * nmax);
        const long nrows = long(nz) * ny * nc;
        #pragma omp parallel for schedule(static)
        for (long p = 0; p < (nrows + 1) / 2; p++) {
            C* w = &work[size_t(omp_get_thread_num()) * 2 * nmax];
            const long r0 = 2 * p, r1 = 2 * p + 1;
            auto rowc = [&](long r) { const int c = r % nc; const long zy = r / nc;
                                      return tmp.data() + size_t(zy) * nxh * nc + c; };
            auto rowr = [&](long r) { const int c = r % nc; const long zy = r / nc;
                                      return buf + size_t(zy) * (nx + 2) * nc + c; };
            r2c_pair(rowr(r0), r1 < nrows ? rowr(r1) : nullptr, nc, rowc(r0),
                     r1 < nrows ? rowc(r1) : nullptr, nc, w, w + nmax);
        }
        #pragma omp parallel for schedule(static)
        for (long line = 0; line < long(nz) * nxh * nc; line++) {
            C* w = &work[size_t(omp_get_thread_num()) * 2 * nmax];
            const int c = line % nc, kx = (line / nc) % nxh, z = line / (nc * nxh);
            C* base = tmp.data() + (size_t(z) * ny * nxh + kx) * nc + c;
            const size_t st = size_t(nxh) * nc;
            for (int y = 0; y < ny; y++) w[y] = base[y * st];
            fy.run(w, w + nmax, -1);
            for (int y = 0; y < ny; y++) base[y * st] = w[y];
        }
        #pragma omp parallel for schedule(static)
        for (long zy = 0; zy < long(nz) * ny; zy++) {
            const int z = zy / ny, ky = zy % ny;
            std::copy(tmp.begin() + (size_t(z) * ny + ky) * nxh * nc,
                      tmp.begin() + (size_t(z) * ny + ky + 1) * nxh * nc,
                      cd + (size_t(ky) * nz + z) * nxh * nc);
        }
        #pragma omp parallel for schedule(static)
        for (long line = 0; line < long(ny) * nxh * nc; line++) {
            C* w = &work[size_t(omp_get_thread_num()) * 2 * nmax];
            const int c = line % nc, kx = (line / nc) % nxh, ky = line / (nc * nxh);
            C* base = cd + (size_t(ky) * nz * nxh + kx) * nc + c;
            const size_t st = size_t(nxh) * nc;
            for (int z = 0; z < nz; z++) w[z] = base[z * st];
            fz.run(w, w + nmax, -1);
            for (int z = 0; z < nz; z++) base[z * st] = w[z];
        }
    }
};

// ------------------------------------------------------------------------ //
// kspace (kspace.cpp:34-167)
// ------------------------------------------------------------------------ //
struct KSpace {
    int nx, ny, nz, nxh;
    bool smooth;
    double dkx, dky, dkz, dk, dk2, kM, kM2;
    int nshells;
    std::vector<double> kx, ky, kz, kshell;
    std::vector<int64_t> nshell;
    KSpace(int nx_, int ny_, int nz_, double dkx_, double dky_, double dkz_, bool smooth_)
        : nx(nx_), ny(ny_), nz(nz_), nxh(nx_ / 2 + 1), smooth(smooth_), dkx(dkx_), dky(dky_), dkz(dkz_) {
        kx.resize(nxh); ky.resize(ny); kz.resize(nz);
        for (int i = 0; i < nxh; i++) kx[i] = i * dkx;
        for (int i = 0; i < ny; i++) ky[i] = (i <= ny / 2) ? dky * i : dky * (i - ny);
        for (int i = 0; i < nz; i++) kz[i] = (i <= nz / 2) ? dkz * i : dkz * (i - nz);
        double kMx, kMy, kMz;
        if (!smooth) {
            kMx = dkx * (int(2 * (nxh - 1) / 3) - 1);
            kMy = dky * (int(ny / 3) - 1);
            kMz = dkz * (int(nz / 3) - 1);
        } else {
            kMx = dkx * (nxh - 2);
            kMy = dky * (int(ny / 2) - 1);
            kMz = dkz * (int(nz / 2) - 1);
        }
        kM = std::max(kMx, std::max(kMy, kMz));
        kM2 = kM * kM;
        dk = std::min(dkx, std::min(dky, dkz));
        dk2 = dk * dk;
        nshells = int(kM / dk) + 2;
        kshell.assign(nshells, 0);
        nshell.assign(nshells, 0);
        for (int y = 0; y < ny; y++)
            for (int z = 0; z < nz; z++)
                for (int x = 0; x < nxh; x++) {
                    double k2 = kx[x] * kx[x] + ky[y] * ky[y] + kz[z] * kz[z];
                    int nxm = x == 0 ? 1 : 2;
                    if (k2 < kM2) {
                        double kn = std::sqrt(k2);
                        kshell[int(kn / dk)] += nxm * kn;
                        nshell[int(kn / dk)] += nxm;
                    }
                }
        for (int s = 0; s < nshells; s++)
            if (nshell[s]) kshell[s] /= nshell[s];
    }
    // smooth filter exp(-36 (k^2/kM^2)^18), table keyed by round(k^2/dk^2) (kspace.cpp:135,479)
    double filter(double k2) const {
        double key = std::round(k2 / dk2);
        (void)key;
        return std::exp(-36.0 * std::pow(k2 / kM2, 18.));
    }
    size_t csize() const { return size_t(ny) * nz * nxh; }
    template <class F> void cloop_k2(F f) const {
        #pragma omp parallel for schedule(static)
        for (long yz = 0; yz < long(ny) * nz; yz++) {
            const int y = yz / nz, z = yz % nz;
            size_t ci = size_t(yz) * nxh;
            for (int x = 0; x < nxh; x++, ci++) {
                double k2 = kx[x] * kx[x] + ky[y] * ky[y] + kz[z] * kz[z];
                f(ci, x, y, z, k2);
            }
        }
    }
};

// ------------------------------------------------------------------------ //
// field-like buffer with the reference's in-place representation
// ------------------------------------------------------------------------ //
template <typename R>
struct Field {
    int nx, ny, nz, nc;
    std::vector<R> data;
    Field(int nx_, int ny_, int nz_, int nc_) : nx(nx_), ny(ny_), nz(nz_), nc(nc_),
        data(size_t(nz_) * ny_ * (nx_ + 2) * nc_, R(0)) {}
    R& cv(size_t ci, int c, int im) { return data[(ci * nc + c) * 2 + im]; }
    R& rv(size_t ri, int c) { return data[ri * nc + c]; }
    void zero() { std::fill(data.begin(), data.end(), R(0)); }
};

template <typename R>
void symmetrize(Field<R>& f) {  // field.cpp:900-978, one rank
    const int nc = f.nc, nxh = f.nx / 2 + 1, ny = f.ny, nz = f.nz;
    R* d = f.data.data();
    auto at = [&](int ky, int kz, int c, int im) -> R& {
        return d[((size_t(ky) * nz + kz) * nxh * nc + c) * 2 + im]; };
    for (int c = 0; c < nc; c++) at(0, 0, c, 1) = 0;
    for (int i = 1; i < nz / 2; i++)
        for (int c = 0; c < nc; c++) {
            at(0, nz - i, c, 0) = at(0, i, c, 0);
            at(0, nz - i, c, 1) = -at(0, i, c, 1);
        }
    std::vector<R> buf(size_t(nz) * nc * 2);
    for (int yy = 1; yy < ny / 2; yy++) {
        for (int i = 0; i < nz; i++)
            for (int c = 0; c < nc; c++)
                for (int im = 0; im < 2; im++) buf[(size_t(i) * nc + c) * 2 + im] = at(yy, i, c, im);
        for (int i = 1; i < nz; i++)
            for (int c = 0; c < nc; c++) {
                at(ny - yy, i, c, 0) = buf[(size_t(nz - i) * nc + c) * 2 + 0];
                at(ny - yy, i, c, 1) = -buf[(size_t(nz - i) * nc + c) * 2 + 1];
            }
        for (int c = 0; c < nc; c++) {
            at(ny - yy, 0, c, 0) = buf[c * 2 + 0];
            at(ny - yy, 0, c, 1) = -buf[c * 2 + 1];
        }
    }
}

// ------------------------------------------------------------------------ //
// vorticity_equation (vorticity_equation.cpp:110-439)
// ------------------------------------------------------------------------ //
// F_FIXED_ENERGY is not a reference forcing: the band 1.5 <= |k| <= 3.5
// forcing of the 1024^3 workload, f = eps u_band / (2 E_band), i.e. the linear
// forcing of the band with the amplitude that makes the injection rate
// sum_band nxmodes Re(u* . f) equal eps (famplitude holds eps).
enum Forcing { F_NONE = 0, F_LINEAR = 1, F_KOLMOGOROV = 2, F_FIXED_ENERGY = 3 };

struct SolverBase {
    virtual ~SolverBase() {}
    virtual void set_cfield(int which, const double* src) = 0;
    virtual void get_cfield(int which, double* dst) = 0;
    virtual void get_rfield(int which, double* dst) = 0;
    virtual void io_projection() = 0;
    virtual void step(double dt) = 0;
    virtual void compute_velocity_of_vorticity() = 0;
    virtual void velocity_rspace() = 0;
    virtual void compute_vorticity_from_velocity() = 0;
    virtual void gradient_rspace() = 0;
    virtual void cospectrum(int a, int b, double* out) = 0;
    virtual void pressure_out(double* out) = 0;
    virtual void acceleration_out(double* out) = 0;
    virtual const KSpace& kspace() const = 0;
    // sample real-space field `which` (1 = velocity, 2 = gradient) at positions
    virtual void interpolate(int which, const double* pos, long np, int n, const double* tab,
                             int deg, int deriv_x, int deriv_y, int deriv_z, double* out) = 0;
    double nu = 0.1, famplitude = 1.0, fk0 = 2.0, fk1 = 4.0;
    int fmode = 1, forcing = F_LINEAR;
    int iteration = 0;
};

template <typename R>
struct Solver : SolverBase {
    KSpace kk;
    Fft3d<R> fft3, fft9;
    Field<R> cvorticity, rvorticity, v1, v2, cvelocity;   // v0 = v3 = cvorticity
    Field<R> rvelocity, grad;                               // particle-side real fields
    Solver(int nx, int ny, int nz, double dkx, double dky, double dkz, bool smooth)
        : kk(nx, ny, nz, dkx, dky, dkz, smooth), fft3(nx, ny, nz, 3), fft9(nx, ny, nz, 9),
          cvorticity(nx, ny, nz, 3), rvorticity(nx, ny, nz, 3), v1(nx, ny, nz, 3),
          v2(nx, ny, nz, 3), cvelocity(nx, ny, nz, 3), rvelocity(nx, ny, nz, 3), grad(nx, ny, nz, 9) {}
    const KSpace& kspace() const override { return kk; }

    Field<R>& sel(int which) {
        switch (which) {
        case 0: return cvorticity;
        case 1: return cvelocity;
        case 2: return v1;
        case 3: return v2;
        case 4: return rvorticity;
        case 5: return rvelocity;
        case 6: return grad;
        }
        throw std::runtime_error("bad field id");
    }
    void set_cfield(int which, const double* src) override {
        Field<R>& f = sel(which);
        const size_t n = kk.csize() * f.nc * 2;
        f.zero();
        for (size_t i = 0; i < n; i++) f.data[i] = R(src[i]);
    }
    void get_cfield(int which, double* dst) override {
        Field<R>& f = sel(which);
        const size_t n = kk.csize() * f.nc * 2;
        for (size_t i = 0; i < n; i++) dst[i] = double(f.data[i]);
    }
    // real field without padding: [z][y][x][c]
    void get_rfield(int which, double* dst) override {
        Field<R>& f = sel(which);
        const int nx = kk.nx, nc = f.nc;
        for (long zy = 0; zy < long(kk.nz) * kk.ny; zy++)
            for (int x = 0; x < nx; x++)
                for (int c = 0; c < nc; c++)
                    dst[(size_t(zy) * nx + x) * nc + c] = double(f.data[(size_t(zy) * (nx + 2) + x) * nc + c]);
    }

    // kspace::low_pass (kspace.cpp:235-253)
    void low_pass(Field<R>& a, double kmax) {
        const double km2 = kmax * kmax;
        kk.cloop_k2([&](size_t ci, int, int, int, double k2) {
            if (k2 >= km2) for (int c = 0; c < a.nc; c++) { a.cv(ci, c, 0) = 0; a.cv(ci, c, 1) = 0; }
        });
    }
    // kspace::dealias (kspace.cpp:461-485)
    void dealias(Field<R>& a) {
        if (!kk.smooth) { low_pass(a, kk.kM); return; }
        kk.cloop_k2([&](size_t ci, int, int, int, double k2) {
            double t = kk.filter(k2);
            for (int c = 0; c < a.nc; c++)
                for (int i = 0; i < 2; i++) a.cv(ci, c, i) = R(a.cv(ci, c, i) * t);
        });
    }
    // kspace::force_divfree (kspace.cpp:487-519)
    void force_divfree(Field<R>& a) {
        kk.cloop_k2([&](size_t ci, int x, int y, int z, double k2) {
            if (k2 > 0) {
                R t0 = R((kk.kx[x] * a.cv(ci, 0, 0) + kk.ky[y] * a.cv(ci, 1, 0) + kk.kz[z] * a.cv(ci, 2, 0)) / k2);
                R t1 = R((kk.kx[x] * a.cv(ci, 0, 1) + kk.ky[y] * a.cv(ci, 1, 1) + kk.kz[z] * a.cv(ci, 2, 1)) / k2);
                R tv[2] = {t0, t1};
                for (int im = 0; im < 2; im++) {
                    a.cv(ci, 0, im) = R(a.cv(ci, 0, im) - tv[im] * kk.kx[x]);
                    a.cv(ci, 1, im) = R(a.cv(ci, 1, im) - tv[im] * kk.ky[y]);
                    a.cv(ci, 2, im) = R(a.cv(ci, 2, im) - tv[im] * kk.kz[z]);
                }
            }
        });
        for (int i = 0; i < 6; i++) a.data[i] = 0;
    }
    // i k x a / k^2 (compute_velocity, vorticity_equation.cpp:206-239)
    void compute_velocity(Field<R>& vort) {
        Field<R>& u = cvelocity;
        kk.cloop_k2([&](size_t ci, int x, int y, int z, double k2) {
            if (k2 <= kk.kM2 && k2 > 0) {
                const double kx = kk.kx[x], ky = kk.ky[y], kz = kk.kz[z];
                u.cv(ci, 0, 0) = R(-(ky * vort.cv(ci, 2, 1) - kz * vort.cv(ci, 1, 1)) / k2);
                u.cv(ci, 0, 1) = R((ky * vort.cv(ci, 2, 0) - kz * vort.cv(ci, 1, 0)) / k2);
                u.cv(ci, 1, 0) = R(-(kz * vort.cv(ci, 0, 1) - kx * vort.cv(ci, 2, 1)) / k2);
                u.cv(ci, 1, 1) = R((kz * vort.cv(ci, 0, 0) - kx * vort.cv(ci, 2, 0)) / k2);
                u.cv(ci, 2, 0) = R(-(kx * vort.cv(ci, 1, 1) - ky * vort.cv(ci, 0, 1)) / k2);
                u.cv(ci, 2, 1) = R((kx * vort.cv(ci, 1, 0) - ky * vort.cv(ci, 0, 0)) / k2);
            } else {
                for (int c = 0; c < 3; c++) { u.cv(ci, c, 0) = 0; u.cv(ci, c, 1) = 0; }
            }
        });
        symmetrize(u);
    }
    void compute_velocity_of_vorticity() override { compute_velocity(cvorticity); }
    // compute_vorticity (vorticity_equation.cpp:171-204): cvorticity = i k x cvelocity
    void compute_vorticity_from_velocity() override {
        Field<R>& u = cvelocity;
        Field<R>& w = cvorticity;
        kk.cloop_k2([&](size_t ci, int x, int y, int z, double k2) {
            if (k2 <= kk.kM2) {
                const double kx = kk.kx[x], ky = kk.ky[y], kz = kk.kz[z];
                w.cv(ci, 0, 0) = R(-(ky * u.cv(ci, 2, 1) - kz * u.cv(ci, 1, 1)));
                w.cv(ci, 0, 1) = R((ky * u.cv(ci, 2, 0) - kz * u.cv(ci, 1, 0)));
                w.cv(ci, 1, 0) = R(-(kz * u.cv(ci, 0, 1) - kx * u.cv(ci, 2, 1)));
                w.cv(ci, 1, 1) = R((kz * u.cv(ci, 0, 0) - kx * u.cv(ci, 2, 0)));
                w.cv(ci, 2, 0) = R(-(kx * u.cv(ci, 1, 1) - ky * u.cv(ci, 0, 1)));
                w.cv(ci, 2, 1) = R((kx * u.cv(ci, 1, 0) - ky * u.cv(ci, 0, 0)));
            } else {
                for (int c = 0; c < 3; c++) { w.cv(ci, c, 0) = 0; w.cv(ci, c, 1) = 0; }
            }
        });
        symmetrize(w);
    }
    // add_forcing (vorticity_equation.cpp:241-288)
    void add_forcing(Field<R>& dst, Field<R>& vort, R factor) {
        if (forcing == F_NONE) return;
        const size_t nxh = kk.nxh, nz = kk.nz;
        if (forcing == F_KOLMOGOROV) {
            size_t ci = size_t(fmode) * nz * nxh;
            dst.cv(ci, 2, 0) = R(dst.cv(ci, 2, 0) - famplitude * factor / 2);
            ci = size_t(kk.ny - fmode) * nz * nxh;
            dst.cv(ci, 2, 0) = R(dst.cv(ci, 2, 0) - famplitude * factor / 2);
            return;
        }
        double amp = famplitude;
        if (forcing == F_FIXED_ENERGY) {
            // E_band = 1/2 sum nxmodes |u|^2 with |u|^2 = |omega|^2 / k^2
            // (serial: a reduction)
            double e = 0.0;
            for (int y = 0; y < kk.ny; y++)
                for (int z = 0; z < int(nz); z++)
                    for (int x = 0; x < int(nxh); x++) {
                        const double k2 = kk.kx[x] * kk.kx[x] + kk.ky[y] * kk.ky[y] + kk.kz[z] * kk.kz[z];
                        const double kn = std::sqrt(k2);
                        if (!(fk0 <= kn && fk1 >= kn && k2 > 0)) continue;
                        const size_t ci = (size_t(y) * nz + z) * nxh + x;
                        double m = 0.0;
                        for (int c = 0; c < 3; c++)
                            for (int i = 0; i < 2; i++) m += double(vort.cv(ci, c, i)) * double(vort.cv(ci, c, i));
                        e += (x == 0 ? 1.0 : 2.0) * m / k2;
                    }
            e *= 0.5;
            amp = e > 0 ? famplitude / (2 * e) : 0.0;
        }
        kk.cloop_k2([&](size_t ci, int x, int y, int z, double) {
            double kn = std::sqrt(kk.kx[x] * kk.kx[x] + kk.ky[y] * kk.ky[y] + kk.kz[z] * kk.kz[z]);
            if (fk0 <= kn && fk1 >= kn)
                for (int c = 0; c < 3; c++)
                    for (int i = 0; i < 2; i++)
                        dst.cv(ci, c, i) = R(dst.cv(ci, c, i) + amp * vort.cv(ci, c, i) * factor);
        });
    }
    // omega_nonlin (vorticity_equation.cpp:290-356)
    void omega_nonlin(Field<R>& src) {
        Field<R>& u = cvelocity;
        compute_velocity(src);
        fft3.c2r(u.data.data());
        rvorticity.data = src.data;
        fft3.c2r(rvorticity.data.data());
        const R npoints = R(double(size_t(kk.nx) * kk.ny * kk.nz));
        const int nx = kk.nx;
        #pragma omp parallel for schedule(static)
        for (long zy = 0; zy < long(kk.nz) * kk.ny; zy++)
            for (int x = 0; x < nx; x++) {
                size_t ri = size_t(zy) * (nx + 2) + x;
                R tmp[3];
                for (int cc = 0; cc < 3; cc++)
                    tmp[cc] = R(u.rv(ri, (cc + 1) % 3) * rvorticity.rv(ri, (cc + 2) % 3) -
                                u.rv(ri, (cc + 2) % 3) * rvorticity.rv(ri, (cc + 1) % 3));
                for (int cc = 0; cc < 3; cc++) u.rv(ri, cc) = tmp[cc] / npoints;
            }
        fft3.r2c(u.data.data());
        dealias(u);
        kk.cloop_k2([&](size_t ci, int x, int y, int z, double) {
            const double kx = kk.kx[x], ky = kk.ky[y], kz = kk.kz[z];
            R t[3][2];
            t[0][0] = R(-(ky * u.cv(ci, 2, 1) - kz * u.cv(ci, 1, 1)));
            t[1][0] = R(-(kz * u.cv(ci, 0, 1) - kx * u.cv(ci, 2, 1)));
            t[2][0] = R(-(kx * u.cv(ci, 1, 1) - ky * u.cv(ci, 0, 1)));
            t[0][1] = R((ky * u.cv(ci, 2, 0) - kz * u.cv(ci, 1, 0)));
            t[1][1] = R((kz * u.cv(ci, 0, 0) - kx * u.cv(ci, 2, 0)));
            t[2][1] = R((kx * u.cv(ci, 1, 0) - ky * u.cv(ci, 0, 0)));
            for (int c = 0; c < 3; c++) for (int i = 0; i < 2; i++) u.cv(ci, c, i) = t[c][i];
        });
        add_forcing(u, src, R(1.0));
        force_divfree(u);
    }
    // step (vorticity_equation.cpp:358-439)
    void step(double dt) override {
        Field<R>& u = cvelocity;
        Field<R>& v0 = cvorticity;
        v1.zero();
        omega_nonlin(v0);
        kk.cloop_k2([&](size_t ci, int, int, int, double k2) {
            if (k2 <= kk.kM2) {
                double f0 = std::exp(-nu * k2 * dt);
                for (int c = 0; c < 3; c++) for (int i = 0; i < 2; i++)
                    v1.cv(ci, c, i) = R((v0.cv(ci, c, i) + dt * u.cv(ci, c, i)) * f0);
            }
        });
        omega_nonlin(v1);
        kk.cloop_k2([&](size_t ci, int, int, int, double k2) {
            if (k2 <= kk.kM2) {
                double f0 = std::exp(-nu * k2 * dt / 2);
                double f1 = std::exp(nu * k2 * dt / 2);
                for (int c = 0; c < 3; c++) for (int i = 0; i < 2; i++)
                    v2.cv(ci, c, i) = R((3 * v0.cv(ci, c, i) * f0 + (v1.cv(ci, c, i) + dt * u.cv(ci, c, i)) * f1) * 0.25);
            }
        });
        omega_nonlin(v2);
        kk.cloop_k2([&](size_t ci, int, int, int, double k2) {
            if (k2 <= kk.kM2) {
                double f0 = std::exp(-nu * k2 * dt * 0.5);
                for (int c = 0; c < 3; c++) for (int i = 0; i < 2; i++)
                    v0.cv(ci, c, i) = R((v0.cv(ci, c, i) * f0 + 2 * (v2.cv(ci, c, i) + dt * u.cv(ci, c, i))) * f0 / 3);
            }
        });
        force_divfree(v0);
        symmetrize(v0);
        iteration++;
    }
    // io_checkpoint(read) projections (vorticity_equation.hpp:107-128; field.cpp:340)
    void io_projection() override {
        symmetrize(cvorticity);
        low_pass(cvorticity, kk.kM);
        force_divfree(cvorticity);
    }
    // NSVEparticles::step first two lines (NSVEparticles.cpp:36-37)
    void velocity_rspace() override {
        compute_velocity(cvorticity);
        rvelocity.data = cvelocity.data;
        fft3.c2r(rvelocity.data.data());
    }
    // compute_gradient<THREE, THREExTHREE> (field.cpp:1042-1101) of cvelocity + ift
    void gradient_rspace() override {
        compute_velocity(cvorticity);
        grad.zero();
        Field<R>& u = cvelocity;
        kk.cloop_k2([&](size_t ci, int x, int y, int z, double k2) {
            if (k2 < kk.kM2) {
                const double kv[3] = {kk.kx[x], kk.ky[y], kk.kz[z]};
                for (int fc = 0; fc < 3; fc++)
                    for (int j = 0; j < 3; j++) {
                        grad.data[((ci * 3 + j) * 3 + fc) * 2 + 0] = R(-kv[j] * u.cv(ci, fc, 1));
                        grad.data[((ci * 3 + j) * 3 + fc) * 2 + 1] = R(kv[j] * u.cv(ci, fc, 0));
                    }
            }
        });
        fft9.c2r(grad.data.data());
    }
    // compute_pressure (vorticity_equation.cpp:441-520) of the real-space
    // velocity in rvelocity (the reference takes `u` in real space):
    // p = -k_i k_j F[u_i u_j] / (npoints k^2), products dealiased, diagonal
    // terms first, then the off-diagonal ones
    void compute_pressure(Field<R>& p) {
        Field<R> prod(kk.nx, kk.ny, kk.nz, 3);
        const int nx = kk.nx;
        const double npoints = double(size_t(kk.nx) * kk.ny * kk.nz);
        for (int pass = 0; pass < 2; pass++) {
            prod.zero();
            #pragma omp parallel for schedule(static)
            for (long zy = 0; zy < long(kk.nz) * kk.ny; zy++)
                for (int x = 0; x < nx; x++) {
                    const size_t ri = size_t(zy) * (nx + 2) + x;
                    for (int cc = 0; cc < 3; cc++)
                        prod.rv(ri, cc) = rvelocity.rv(ri, cc) * rvelocity.rv(ri, pass == 0 ? cc : (cc + 1) % 3);
                }
            fft3.r2c(prod.data.data());
            dealias(prod);
            kk.cloop_k2([&](size_t ci, int x, int y, int z, double k2) {
                const double kx = kk.kx[x], ky = kk.ky[y], kz = kk.kz[z];
                if (k2 <= kk.kM2 && k2 > 0) {
                    for (int i = 0; i < 2; i++) {
                        if (pass == 0) {
                            p.cv(ci, 0, i) = R(-(kx * kx * prod.cv(ci, 0, i) + ky * ky * prod.cv(ci, 1, i) +
                                                 kz * kz * prod.cv(ci, 2, i)));
                        } else {
                            p.cv(ci, 0, i) = R(p.cv(ci, 0, i) - 2 * (kx * ky * prod.cv(ci, 0, i) +
                                                                     ky * kz * prod.cv(ci, 1, i) +
                                                                     kz * kx * prod.cv(ci, 2, i)));
                            p.cv(ci, 0, i) = R(p.cv(ci, 0, i) / (npoints * k2));
                        }
                    }
                } else if (pass == 0) {
                    p.cv(ci, 0, 0) = 0;
                    p.cv(ci, 0, 1) = 0;
                }
            });
        }
    }
    // compute_Lagrangian_acceleration (vorticity_equation.cpp:522-578):
    // a = -nu k^2 u (+ famplitude u in the band for linear forcing) - i k p
    void lagrangian_acceleration(Field<R>& acc) {
        velocity_rspace();
        Field<R> p(kk.nx, kk.ny, kk.nz, 1);
        compute_pressure(p);
        compute_velocity(cvorticity);
        acc.zero();
        Field<R>& u = cvelocity;
        kk.cloop_k2([&](size_t ci, int x, int y, int z, double k2) {
            if (k2 <= kk.kM2) {
                const double kv[3] = {kk.kx[x], kk.ky[y], kk.kz[z]};
                for (int c = 0; c < 3; c++)
                    for (int i = 0; i < 2; i++) acc.cv(ci, c, i) = R(-nu * k2 * u.cv(ci, c, i));
                if (forcing == F_LINEAR) {
                    const double kn = std::sqrt(k2);
                    if (fk0 <= kn && fk1 >= kn)
                        for (int c = 0; c < 3; c++)
                            for (int i = 0; i < 2; i++) acc.cv(ci, c, i) = R(acc.cv(ci, c, i) + famplitude * u.cv(ci, c, i));
                }
                for (int c = 0; c < 3; c++) acc.cv(ci, c, 0) = R(acc.cv(ci, c, 0) + kv[c] * p.cv(ci, 0, 1));
                for (int c = 0; c < 3; c++) acc.cv(ci, c, 1) = R(acc.cv(ci, c, 1) - kv[c] * p.cv(ci, 0, 0));
            }
        });
    }
    void pressure_out(double* out) override {
        velocity_rspace();
        Field<R> p(kk.nx, kk.ny, kk.nz, 1);
        compute_pressure(p);
        for (size_t i = 0; i < kk.csize() * 2; i++) out[i] = double(p.data[i]);
    }
    void acceleration_out(double* out) override {
        Field<R> a(kk.nx, kk.ny, kk.nz, 3);
        lagrangian_acceleration(a);
        for (size_t i = 0; i < kk.csize() * 6; i++) out[i] = double(a.data[i]);
    }

    // kspace::cospectrum (kspace.cpp:521-598) of two 3-component fields
    void cospectrum(int a, int b, double* out) override {
        Field<R>& fa = sel(a);
        Field<R>& fb = sel(b);
        const int ns = kk.nshells;
        std::fill(out, out + ns * 9, 0.0);
        const int nth = omp_get_max_threads();
        std::vector<double> part(size_t(nth) * ns * 9, 0.0);
        #pragma omp parallel
        {
            double* sp = &part[size_t(omp_get_thread_num()) * ns * 9];
            #pragma omp for schedule(static)
            for (long yz = 0; yz < long(kk.ny) * kk.nz; yz++) {
                const int y = yz / kk.nz, z = yz % kk.nz;
                for (int x = 0; x < kk.nxh; x++) {
                    size_t ci = size_t(yz) * kk.nxh + x;
                    double k2 = kk.kx[x] * kk.kx[x] + kk.ky[y] * kk.ky[y] + kk.kz[z] * kk.kz[z];
                    if (k2 <= kk.kM2) {
                        int nxm = x == 0 ? 1 : 2;
                        int s = int(std::sqrt(k2) / kk.dk) * 9;
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++)
                                sp[s + i * 3 + j] += nxm * (double(fa.cv(ci, i, 0)) * double(fb.cv(ci, j, 0)) +
                                                            double(fa.cv(ci, i, 1)) * double(fb.cv(ci, j, 1)));
                    }
                }
            }
        }
        for (int t = 0; t < nth; t++)
            for (int i = 0; i < ns * 9; i++) out[i] += part[size_t(t) * ns * 9 + i];
    }

    // particles_field_computer::apply_computation (particles_field_computer.hpp:65-157), one rank
    void interpolate(int which, const double* pos, long np, int n, const double* tab, int deg,
                     int deriv_x, int deriv_y, int deriv_z, double* out) override {
        Field<R>& f = (which == 1) ? rvelocity : grad;
        const int nc = f.nc;
        const int N[3] = {kk.nx, kk.ny, kk.nz};
        const double dkv[3] = {kk.dkx, kk.dky, kk.dkz};
        double width[3], stepw[3];
        for (int d = 0; d < 3; d++) {
            width[d] = 4 * std::acos(0.0) / dkv[d];
            stepw[d] = width[d] / double(N[d]);
        }
        const int W = 2 * n + 2;
        const int der[3] = {deriv_x, deriv_y, deriv_z};
        #pragma omp parallel for schedule(static)
        for (long p = 0; p < np; p++) {
            double b[3][32];
            int cell[3];
            for (int d = 0; d < 3; d++) {
                const double s = pos[p * 3 + d] - 0.0;
                const double rep = std::floor(s / width[d]);
                const double inbox = s - rep * width[d];
                const double ci = std::floor(inbox / stepw[d]);
                const double xr = (inbox - ci * stepw[d]) / stepw[d];
                const double* tb = tab + size_t(der[d]) * W * (deg + 1);
                for (int j = 0; j < W; j++) {
                    double acc = 0;
                    for (int q = deg; q >= 0; q--) acc = acc * xr + tb[j * (deg + 1) + q];
                    b[d][j] = acc;
                }
                const int lvl = int(std::floor(s / stepw[d]));
                cell[d] = ((lvl % N[d]) + N[d]) % N[d];
            }
            double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int jz = 0; jz < W; jz++) {
                const int zz = ((cell[2] - n + jz) % N[2] + N[2]) % N[2];
                for (int jx = 0; jx < W; jx++) {
                    const int xx = ((cell[0] - n + jx) % N[0] + N[0]) % N[0];
                    for (int jy = 0; jy < W; jy++) {
                        const int yy = ((cell[1] - n + jy) % N[1] + N[1]) % N[1];
                        const double coef = b[2][jz] * b[1][jy] * b[0][jx];
                        const size_t ri = (size_t(zz) * kk.ny + yy) * (kk.nx + 2) + xx;
                        for (int c = 0; c < nc; c++) acc[c] += double(f.data[ri * nc + c]) * coef;
                    }
                }
            }
            for (int c = 0; c < nc; c++) out[p * nc + c] = acc[c];
        }
    }
};

// ------------------------------------------------------------------------ //
// interpolation weight tables from the definition (numeric)
// ------------------------------------------------------------------------ //
typedef std::vector<double> Poly;  // lowest order first

static Poly pmul(const Poly& a, const Poly& b) {
    Poly r(a.size() + b.size() - 1, 0.0);
    for (size_t i = 0; i < a.size(); i++)
        for (size_t j = 0; j < b.size(); j++) r[i + j] += a[i] * b[j];
    return r;
}
static Poly pder(const Poly& a) {
    if (a.size() <= 1) return Poly(1, 0.0);
    Poly r(a.size() - 1);
    for (size_t i = 1; i < a.size(); i++) r[i - 1] = a[i] * double(i);
    return r;
}
static double peval(const Poly& a, double x) {
    double acc = 0;
    for (size_t i = a.size(); i-- > 0;) acc = acc * x + a[i];
    return acc;
}
static Poly lagrange(const std::vector<int>& nodes, int node) {
    Poly r(1, 1.0);
    for (int o : nodes)
        if (o != node) {
            Poly f = {-double(o) / double(node - o), 1.0 / double(node - o)};
            r = pmul(r, f);
        }
    return r;
}
static std::vector<double> solve(std::vector<double> A, std::vector<double> b, int n) {
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++) if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
        for (int k = 0; k < n; k++) std::swap(A[c * n + k], A[piv * n + k]);
        std::swap(b[c], b[piv]);
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            double f = A[r * n + c] / A[c * n + c];
            for (int k = 0; k < n; k++) A[r * n + k] -= f * A[c * n + k];
            b[r] -= f * b[c];
        }
    }
    for (int c = 0; c < n; c++) b[c] /= A[c * n + c];
    return b;
}

// table [deriv 0..2][W][deg+1]; returns degree
static int build_table(int n, int m, std::vector<double>& tab) {
    const int W = 2 * n + 2;
    int deg;
    std::vector<Poly> w(W);
    if (m == 0) {
        deg = 2 * n + 1;
        std::vector<int> nodes;
        for (int i = -n; i <= n + 1; i++) nodes.push_back(i);
        for (int j = 0; j < W; j++) w[j] = lagrange(nodes, j - n);
    } else {
        deg = 2 * m + 1;
        std::vector<int> n0, n1;
        for (int i = -n; i <= n; i++) n0.push_back(i);
        for (int i = 1 - n; i <= n + 1; i++) n1.push_back(i);
        const int K = deg + 1;
        std::vector<double> A(K * K, 0.0);
        // rows: value/derivs at 0 (k = 0..m), then at 1
        for (int k = 0; k <= m; k++)
            for (int q = 0; q <= deg; q++) {
                // d^k/dx^k x^q at 0 and 1
                double c = 1.0;
                for (int t = 0; t < k; t++) c *= double(q - t);
                A[k * K + q] = (q == k) ? c : 0.0;
                A[(m + 1 + k) * K + q] = (q >= k) ? c : 0.0;
            }
        for (int j = 0; j < W; j++) {
            const int node = j - n;
            Poly l0 = (node >= -n && node <= n) ? lagrange(n0, node) : Poly(1, 0.0);
            Poly l1 = (node >= 1 - n && node <= n + 1) ? lagrange(n1, node) : Poly(1, 0.0);
            std::vector<double> rhs(K);
            Poly d0 = l0, d1 = l1;
            for (int k = 0; k <= m; k++) {
                rhs[k] = peval(d0, 0.0);
                rhs[m + 1 + k] = peval(d1, 1.0);
                d0 = pder(d0);
                d1 = pder(d1);
            }
            w[j] = solve(A, rhs, K);
        }
    }
    tab.assign(size_t(3) * W * (deg + 1), 0.0);
    for (int d = 0; d < 3; d++)
        for (int j = 0; j < W; j++) {
            Poly p = w[j];
            for (int t = 0; t < d; t++) p = pder(p);
            for (size_t q = 0; q < p.size() && q <= size_t(deg); q++)
                tab[(size_t(d) * W + j) * (deg + 1) + q] = p[q];
        }
    return deg;
}

// ------------------------------------------------------------------------ //
// particles_system (particles_system.hpp:173-217), one rank, order kept
// ------------------------------------------------------------------------ //
struct Particles {
    SolverBase* s;
    long np;
    int n, m, nsteps, step_idx, deg;
    std::vector<double> tab;
    std::vector<double> pos;
    std::vector<std::vector<double>> rhs;
    Particles(SolverBase* s_, long np_, int n_, int m_, int nsteps_, int step_idx_)
        : s(s_), np(np_), n(n_), m(m_), nsteps(nsteps_), step_idx(step_idx_) {
        if (nsteps > 6) throw std::runtime_error("Adams-Bashforth order above 6");
        deg = build_table(n, m, tab);
        pos.assign(np * 3, 0.0);
        rhs.assign(nsteps, std::vector<double>(np * 3, 0.0));
    }
    // particles_adams_bashforth::move_particles (particles_adams_bashforth.hpp:41-131)
    void move(double dt) {
        const int k = std::min(step_idx, nsteps);
        const long nv = np * 3;
        const double* r0 = k > 0 ? rhs[0].data() : nullptr;
        const double* r1 = k > 1 ? rhs[1].data() : nullptr;
        const double* r2 = k > 2 ? rhs[2].data() : nullptr;
        const double* r3 = k > 3 ? rhs[3].data() : nullptr;
        const double* r4 = k > 4 ? rhs[4].data() : nullptr;
        const double* r5 = k > 5 ? rhs[5].data() : nullptr;
        #pragma omp parallel for schedule(static)
        for (long i = 0; i < nv; i++) {
            switch (k) {
            case 1: pos[i] += dt * r0[i]; break;
            case 2: pos[i] += dt * (3. * r0[i] - r1[i]) / 2.; break;
            case 3: pos[i] += dt * (23. * r0[i] - 16. * r1[i] + 5. * r2[i]) / 12.; break;
            case 4: pos[i] += dt * (55. * r0[i] - 59. * r1[i] + 37. * r2[i] - 9. * r3[i]) / 24.; break;
            case 5: pos[i] += dt * (1901. * r0[i] - 2774. * r1[i] + 2616. * r2[i] - 1274. * r3[i] + 251. * r4[i]) / 720.; break;
            case 6: pos[i] += dt * (4277. * r0[i] - 7923. * r1[i] + 9982. * r2[i] - 7298. * r3[i] + 2877. * r4[i] - 475. * r5[i]) / 1440.; break;
            }
        }
    }
    void complete_loop(double dt) {
        s->interpolate(1, pos.data(), np, n, tab.data(), deg, 0, 0, 0, rhs[0].data());
        move(dt);
        step_idx += 1;
        // shift_rhs_vectors
        std::vector<double> last = std::move(rhs.back());
        for (int i = nsteps - 1; i > 0; i--) rhs[i] = std::move(rhs[i - 1]);
        rhs[0] = std::move(last);
        std::fill(rhs[0].begin(), rhs[0].end(), 0.0);
    }
};

}  // namespace orc

using namespace orc;

// ------------------------------------------------------------------------ //
// field::compute_rspace_stats (field.cpp:661-850), one rank: data [npts][nc]
// (the reference's real values), nvals = 4 for 3 components (the magnitude
// is value 3), else nc.  moments [10][nvals]: 0 = min, 9 = max (both start
// at 0, the magnitude's min at its estimate), 1..8 = mean of value^n;
// hist [nbins][nvals]: bins of width 2 max_estimate / nbins from
// -max_estimate (the magnitude's from 0, half the width)
// ------------------------------------------------------------------------ //
static void rspace_stats(const double* d, long npts, int nc, int nbins, const double* maxest, double* moments,
                         int64_t* hist) {
    const int nm = 10;
    const int nvals = nc == 3 ? 4 : nc;
    std::vector<double> mom(size_t(nm) * nvals, 0.0), binsize(nvals), val(nvals), pw(nvals);
    if (nvals == 4) mom[3] = maxest[3];
    std::fill(hist, hist + size_t(nbins) * nvals, int64_t(0));
    for (int i = 0; i < nvals; i++) binsize[i] = 2 * maxest[i] / nbins;
    for (long p = 0; p < npts; p++) {
        std::fill(pw.begin(), pw.end(), 1.0);
        if (nvals == 4) val[3] = 0.0;
        for (int i = 0; i < nc; i++) {
            val[i] = d[size_t(p) * nc + i];
            if (nvals == 4) val[3] += val[i] * val[i];
        }
        if (nvals == 4) {
            val[3] = std::sqrt(val[3]);
            if (val[3] < mom[3]) mom[3] = val[3];
            if (val[3] > mom[9 * nvals + 3]) mom[9 * nvals + 3] = val[3];
            const int bin = int(std::floor(val[3] * 2 / binsize[3]));
            if (bin >= 0 && bin < nbins) hist[size_t(bin) * nvals + 3]++;
        }
        for (int i = 0; i < nc; i++) {
            if (val[i] < mom[i]) mom[i] = val[i];
            if (val[i] > mom[(nm - 1) * nvals + i]) mom[(nm - 1) * nvals + i] = val[i];
            const int bin = int(std::floor((val[i] + maxest[i]) / binsize[i]));
            if (bin >= 0 && bin < nbins) hist[size_t(bin) * nvals + i]++;
        }
        for (int n = 1; n < nm - 1; n++)
            for (int i = 0; i < nvals; i++) mom[size_t(n) * nvals + i] += (pw[i] = val[i] * pw[i]);
    }
    for (int n = 1; n < nm - 1; n++)
        for (int i = 0; i < nvals; i++) mom[size_t(n) * nvals + i] /= double(npts);
    std::copy(mom.begin(), mom.end(), moments);
}

extern "C" {

void* orc_new(int nx, int ny, int nz, double dkx, double dky, double dkz, int fp32, int smooth) {
    if (fp32) return static_cast<SolverBase*>(new Solver<float>(nx, ny, nz, dkx, dky, dkz, smooth != 0));
    return static_cast<SolverBase*>(new Solver<double>(nx, ny, nz, dkx, dky, dkz, smooth != 0));
}
void orc_free(void* h) { delete static_cast<SolverBase*>(h); }
void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_get_threads() { return omp_get_max_threads(); }
void orc_set_physics(void* h, double nu, int forcing, double famp, double fk0, double fk1, int fmode) {
    SolverBase* s = static_cast<SolverBase*>(h);
    s->nu = nu; s->forcing = forcing; s->famplitude = famp; s->fk0 = fk0; s->fk1 = fk1; s->fmode = fmode;
}
void orc_kspace_info(void* h, double* out /* kM, kM2, dk, nshells */) {
    const KSpace& k = static_cast<SolverBase*>(h)->kspace();
    out[0] = k.kM; out[1] = k.kM2; out[2] = k.dk; out[3] = k.nshells;
}
void orc_kshell(void* h, double* kshell, int64_t* nshell) {
    const KSpace& k = static_cast<SolverBase*>(h)->kspace();
    std::copy(k.kshell.begin(), k.kshell.end(), kshell);
    std::copy(k.nshell.begin(), k.nshell.end(), nshell);
}
void orc_set_cfield(void* h, int which, const double* src) { static_cast<SolverBase*>(h)->set_cfield(which, src); }
void orc_get_cfield(void* h, int which, double* dst) { static_cast<SolverBase*>(h)->get_cfield(which, dst); }
void orc_get_rfield(void* h, int which, double* dst) { static_cast<SolverBase*>(h)->get_rfield(which, dst); }
void orc_io_projection(void* h) { static_cast<SolverBase*>(h)->io_projection(); }
void orc_step(void* h, double dt) { static_cast<SolverBase*>(h)->step(dt); }
void orc_compute_velocity(void* h) { static_cast<SolverBase*>(h)->compute_velocity_of_vorticity(); }
void orc_compute_vorticity(void* h) { static_cast<SolverBase*>(h)->compute_vorticity_from_velocity(); }
void orc_velocity_rspace(void* h) { static_cast<SolverBase*>(h)->velocity_rspace(); }
void orc_gradient_rspace(void* h) { static_cast<SolverBase*>(h)->gradient_rspace(); }
void orc_cospectrum(void* h, int a, int b, double* out) { static_cast<SolverBase*>(h)->cospectrum(a, b, out); }
void orc_pressure(void* h, double* out) { static_cast<SolverBase*>(h)->pressure_out(out); }
void orc_lagrangian_acceleration(void* h, double* out) { static_cast<SolverBase*>(h)->acceleration_out(out); }
void orc_rspace_stats(const double* data, long npts, int nc, int nbins, const double* maxest, double* moments,
                      int64_t* hist) {
    rspace_stats(data, npts, nc, nbins, maxest, moments, hist);
}

// 3D FFT on unpadded real [nz][ny][nx][nc] <-> complex [ny][nz][nx/2+1][nc]
void orc_fft3d(int nx, int ny, int nz, int nc, int fp32, int inverse, const double* in, double* out) {
    const size_t nxh = nx / 2 + 1;
    auto go = [&](auto tag) {
        typedef decltype(tag) T;
        Fft3d<T> f(nx, ny, nz, nc);
        std::vector<T> buf(f.rsize(), T(0));
        if (!inverse) {
            for (size_t zy = 0; zy < size_t(nz) * ny; zy++)
                for (int x = 0; x < nx; x++)
                    for (int c = 0; c < nc; c++) buf[(zy * (nx + 2) + x) * nc + c] = T(in[(zy * nx + x) * nc + c]);
            f.r2c(buf.data());
            for (size_t i = 0; i < size_t(ny) * nz * nxh * nc * 2; i++) out[i] = double(buf[i]);
        } else {
            for (size_t i = 0; i < size_t(ny) * nz * nxh * nc * 2; i++) buf[i] = T(in[i]);
            f.c2r(buf.data());
            for (size_t zy = 0; zy < size_t(nz) * ny; zy++)
                for (int x = 0; x < nx; x++)
                    for (int c = 0; c < nc; c++) out[(zy * nx + x) * nc + c] = double(buf[(zy * (nx + 2) + x) * nc + c]);
        }
    };
    if (fp32) go(float()); else go(double());
}

// field::symmetrize (field.cpp:900-978) of a standalone complex field
// [ny][nz][nx/2+1][nc] {re, im}, in place.
void orc_symmetrize(int nx, int ny, int nz, int nc, double* data) {
    Field<double> f(nx, ny, nz, nc);
    const size_t cs = size_t(ny) * nz * (nx / 2 + 1) * nc * 2;
    std::copy(data, data + cs, f.data.begin());
    symmetrize(f);
    std::copy(f.data.begin(), f.data.begin() + cs, data);
}

int orc_interp_table(int n, int m, double* out /* may be null */) {
    std::vector<double> tab;
    int deg = build_table(n, m, tab);
    if (out) std::copy(tab.begin(), tab.end(), out);
    return deg;
}

void* orc_part_new(void* solver, long np, int n, int m, int nsteps, int step_idx) {
    return new Particles(static_cast<SolverBase*>(solver), np, n, m, nsteps, step_idx);
}
void orc_part_free(void* p) { delete static_cast<Particles*>(p); }
void orc_part_set(void* h, const double* pos, const double* rhs) {
    Particles* p = static_cast<Particles*>(h);
    std::copy(pos, pos + p->np * 3, p->pos.begin());
    for (int i = 0; i < p->nsteps; i++) std::copy(rhs + size_t(i) * p->np * 3, rhs + size_t(i + 1) * p->np * 3, p->rhs[i].begin());
}
void orc_part_get(void* h, double* pos, double* rhs, int* step_idx) {
    Particles* p = static_cast<Particles*>(h);
    std::copy(p->pos.begin(), p->pos.end(), pos);
    for (int i = 0; i < p->nsteps; i++) std::copy(p->rhs[i].begin(), p->rhs[i].end(), rhs + size_t(i) * p->np * 3);
    *step_idx = p->step_idx;
}
void orc_part_complete_loop(void* h, double dt) { static_cast<Particles*>(h)->complete_loop(dt); }
// sample velocity (which=1, 3 comps) or gradient (which=2, 9 comps) at current positions
void orc_part_sample(void* h, int which, double* out) {
    Particles* p = static_cast<Particles*>(h);
    p->s->interpolate(which, p->pos.data(), p->np, p->n, p->tab.data(), p->deg, 0, 0, 0, out);
}
// NSVEparticles::step (NSVEparticles.cpp:33-41)
void orc_nsve_particles_step(void* solver, void* parts, double dt) {
    SolverBase* s = static_cast<SolverBase*>(solver);
    s->velocity_rspace();
    if (parts) static_cast<Particles*>(parts)->complete_loop(dt);
    s->step(dt);
}

}  // extern "C"
