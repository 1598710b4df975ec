// C ABI implementation of the MI355X backend (see include/bfpsh.h).
#include <cstdlib>
#include <mutex>

#include "ctx.hpp"
#include "interp_tables.h"

namespace {

const OpsTable* ops_for(const bfpsh_ctx* ctx) {
    switch (ctx->n) {
#define X(N) case N: return ctx->prec == BFPSH_FP32 ? bfpsh_ops_f32_##N() : bfpsh_ops_f64_##N();
        BFPSH_FOR_SIZES(X)
#undef X
    }
    return nullptr;
}

bool supported_size(int n) {
    switch (n) {
#define X(N) case N: return true;
        BFPSH_FOR_SIZES(X)
#undef X
    }
    return false;
}

template <typename T>
void build_twiddles(std::vector<unsigned char>& out, int n) {
    out.resize(size_t(n) * 2 * sizeof(T));
    T* p = reinterpret_cast<T*>(out.data());
    for (int t = 0; t < n; t++) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * t / n;
        p[2 * t] = T(std::cos(a));
        p[2 * t + 1] = T(std::sin(a));
    }
}

}  // namespace

// host [ky][kz][kx=N/2+1][nc]{re,im} <-> device planar [c][ky][kz][kx<N/2]
// (ky rows row_stride(n) complex apart, planes plane_size(n) apart)
template <typename T>
static void host_to_planar(const T* h, T* p, int n, int nc = 3) {
    const int nxh = n / 2 + 1, nxs = n / 2;
    const size_t ps = plane_size(n);
    for (size_t yz = 0; yz < size_t(n) * n; yz++)
        for (int x = 0; x < nxs; x++)
            for (int c = 0; c < nc; c++) {
                const size_t d = c * ps + sphys(n, yz * nxs + x);
                p[d * 2 + 0] = h[((yz * nxh + x) * nc + c) * 2 + 0];
                p[d * 2 + 1] = h[((yz * nxh + x) * nc + c) * 2 + 1];
            }
}
template <typename T>
static void planar_to_host(const T* p, T* h, int n, int nc = 3) {
    const int nxh = n / 2 + 1, nxs = n / 2;
    const size_t ps = plane_size(n);
    for (size_t yz = 0; yz < size_t(n) * n; yz++)
        for (int x = 0; x < nxh; x++)
            for (int c = 0; c < nc; c++) {
                T re = 0, im = 0;
                if (x < nxs) {
                    const size_t d = c * ps + sphys(n, yz * nxs + x);
                    re = p[d * 2 + 0];
                    im = p[d * 2 + 1];
                }
                h[((yz * nxh + x) * nc + c) * 2 + 0] = re;
                h[((yz * nxh + x) * nc + c) * 2 + 1] = im;
            }
}
// host real [z][y][x][nc] <-> device planar [c][y][z][x] (real rows y
// 2 row_stride(n) reals apart, planes 2 plane_size(n) apart)
template <typename T>
static void host_real_to_planar(const T* h, T* p, int n, int nc) {
    const size_t ps = 2 * plane_size(n), ys = 2 * row_stride(n);
    for (size_t y = 0; y < size_t(n); y++)
        for (size_t z = 0; z < size_t(n); z++)
            for (size_t x = 0; x < size_t(n); x++)
                for (int c = 0; c < nc; c++) p[c * ps + y * ys + z * n + x] = h[((z * n + y) * n + x) * nc + c];
}
template <typename T>
static void planar_to_host_real(const T* p, T* h, int n, int nc) {
    const size_t ps = 2 * plane_size(n), ys = 2 * row_stride(n);
    for (size_t y = 0; y < size_t(n); y++)
        for (size_t z = 0; z < size_t(n); z++)
            for (size_t x = 0; x < size_t(n); x++)
                for (int c = 0; c < nc; c++) h[((z * n + y) * n + x) * nc + c] = p[c * ps + y * ys + z * n + x];
}

#define CTX_CHECK()                                  \
    do {                                             \
        if (!ctx) return BFPSH_EINVAL;               \
        hipError_t e_ = hipSetDevice(ctx->device);   \
        if (e_ != hipSuccess) return fail(ctx, BFPSH_EHIP, hipGetErrorString(e_)); \
    } while (0)

#define PREC_CALL(fn, ...) (ops_for(ctx)->fn(__VA_ARGS__))

// ------------------------------------------------------------ fixed scratch
// Shell sums of the allocation-free cospectrum (bfpsh_field_cospectrum_noalloc)
// live in this array, not in a per-call allocation.  Each GPU the library
// runs on holds its own copy until the process ends; host code reaches a
// copy only through the address hipGetSymbolAddress returns for its device
// (resolved once per device when the first context is created there).
// Every call on a GPU shares that GPU's copy: calls that overlap on one GPU
// race.
constexpr int kScratchMaxShells = 1024;
static __device__ __align__(256) double scratch[kScratchMaxShells * 9];
constexpr int kScratchMaxDevices = 64;
static std::mutex scratch_mu;
static double* scratch_dev[kScratchMaxDevices];  // per device: its copy of `scratch`

// resolve this device's copy (the device is current); 0 on success
static int scratch_resolve(bfpsh_ctx* ctx, int device) {
    if (device < 0 || device >= kScratchMaxDevices) return fail(ctx, BFPSH_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (scratch_dev[device]) return BFPSH_OK;
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(ctx, BFPSH_EHIP, "no device address for the cospectrum scratch");
    scratch_dev[device] = static_cast<double*>(p);
    return BFPSH_OK;
}

static __global__ void k_fill_zero_d(double* __restrict__ p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.0;
}

extern "C" {

int bfpsh_create(bfpsh_ctx** out, int nx, int ny, int nz, double dkx, double dky, double dkz, int precision,
                 int dealias, int device) {
    return create_ctx(out, nx, ny, nz, dkx, dky, dkz, precision, dealias, device, nullptr, nullptr);
}

}  // extern "C"

// A context; with `stream` it runs on that (shared) stream, with `slab` its
// v0 | v1 | v2 | work (6 planes) | ru buffers are the 18 planes at `slab`.
int create_ctx(bfpsh_ctx** out, int nx, int ny, int nz, double dkx, double dky, double dkz, int precision,
               int dealias, int device, hipStream_t stream, void* slab) {
    if (!out) return BFPSH_EINVAL;
    *out = nullptr;
    bfpsh_ctx* ctx = new bfpsh_ctx();
    if (nx != ny || ny != nz || !supported_size(nx)) {
        ctx->err = "unsupported grid (cubic N in {16,32,48,64,128,256,512,1024} required)";
        *out = ctx;
        return BFPSH_EINVAL;
    }
    if (precision != BFPSH_FP32 && precision != BFPSH_FP64) { *out = ctx; return fail(ctx, BFPSH_EINVAL, "bad precision"); }
    if (dealias != BFPSH_SMOOTH && dealias != BFPSH_TWO_THIRDS) { *out = ctx; return fail(ctx, BFPSH_EINVAL, "bad dealias type"); }
    if (!(dkx > 0 && dky > 0 && dkz > 0)) { *out = ctx; return fail(ctx, BFPSH_EINVAL, "dk must be positive"); }
    *out = ctx;
    ctx->n = nx;
    ctx->prec = precision;
    ctx->dealias = dealias;
    ctx->device = device;
    ctx->dkx = dkx;
    ctx->dky = dky;
    ctx->dkz = dkz;
    ctx->es = precision == BFPSH_FP32 ? 4 : 8;
    HIPCHK(hipSetDevice(device));
    {
        int r = scratch_resolve(ctx, device);
        if (r) return r;
    }
    if (stream) {
        ctx->stream = stream;
        ctx->own_stream = false;
    } else {
        HIPCHK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    }
    // kspace geometry (bfps/cpp/kspace.cpp:56-166), cubic grid
    const int n = nx, nxh = n / 2 + 1;
    double kMx, kMy, kMz;
    if (dealias == BFPSH_TWO_THIRDS) {
        kMx = dkx * (int(2 * (nxh - 1) / 3) - 1);
        kMy = dky * (int(n / 3) - 1);
        kMz = dkz * (int(n / 3) - 1);
    } else {
        kMx = dkx * (nxh - 2);
        kMy = dky * (int(n / 2) - 1);
        kMz = dkz * (int(n / 2) - 1);
    }
    ctx->kM = std::max(kMx, std::max(kMy, kMz));
    ctx->kM2 = ctx->kM * ctx->kM;
    ctx->dk = std::min(dkx, std::min(dky, dkz));
    ctx->nshells = int(ctx->kM / ctx->dk) + 2;
    ctx->kshell.assign(ctx->nshells, 0.0);
    ctx->nshell.assign(ctx->nshells, 0);
    for (int y = 0; y < n; y++)
        for (int z = 0; z < n; z++)
            for (int x = 0; x < nxh; x++) {
                const double kx = x * dkx;
                const double ky = (y <= n / 2 ? y : y - n) * dky;
                const double kz = (z <= n / 2 ? z : z - n) * dkz;
                const double k2 = kx * kx + ky * ky + kz * kz;
                if (k2 < ctx->kM2) {
                    const double kn = std::sqrt(k2);
                    ctx->kshell[int(kn / ctx->dk)] += (x == 0 ? 1 : 2) * kn;
                    ctx->nshell[int(kn / ctx->dk)] += (x == 0 ? 1 : 2);
                }
            }
    for (int s = 0; s < ctx->nshells; s++)
        if (ctx->nshell[s]) ctx->kshell[s] /= ctx->nshell[s];
    // device state
    ctx->ps = plane_size(n);
    const size_t vbytes = 3 * ctx->ps * 2 * ctx->es;
    if (slab) {
        // 18 planes: v0 (3), v1 (3), v2 (3), work (6), real velocity (3); the
        // owner clears the slab
        char* b = (char*)slab;
        const size_t pb = ctx->ps * 2 * ctx->es;
        ctx->slab = slab;
        ctx->v0 = b;
        ctx->v1 = b + 3 * pb;
        ctx->v2 = b + 6 * pb;
        ctx->work = b + 9 * pb;
        ctx->work_planes = 6;
        ctx->ru = b + 15 * pb;
    } else {
        for (void** p : {&ctx->v0, &ctx->v1, &ctx->v2}) {
            HIPCHK(hipMalloc(p, vbytes));
            HIPCHK(hipMemsetAsync(*p, 0, vbytes, ctx->stream));
            ctx->bytes += vbytes;
        }
    }
    std::vector<unsigned char> tw;
    if (precision == BFPSH_FP32) build_twiddles<float>(tw, n); else build_twiddles<double>(tw, n);
    HIPCHK(hipMalloc(&ctx->tw, tw.size()));
    HIPCHK(hipMemcpyAsync(ctx->tw, tw.data(), tw.size(), hipMemcpyHostToDevice, ctx->stream));
    ctx->bytes += tw.size();
    int r = ensure_work(ctx, 6);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return BFPSH_OK;
}

extern "C" {

int bfpsh_destroy(bfpsh_ctx* ctx) {
    if (!ctx) return BFPSH_OK;
    if (ctx->stream) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (void* p : {ctx->v0, ctx->v1, ctx->v2, ctx->work, ctx->ru, ctx->rgrad, ctx->tw,
                        (void*)ctx->d_pos, (void*)ctx->d_rhs, (void*)ctx->d_tab, (void*)ctx->d_famp,
                        (void*)ctx->d_rank, (void*)ctx->d_samp})
            if (p && !in_slab(ctx, p)) (void)hipFree(p);
        for (void* p : ctx->fields)
            if (p) (void)hipFree(p);
        bfpsh_part_release(ctx);
        resolve_profile(ctx);
        for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return BFPSH_OK;
}

const char* bfpsh_last_error(const bfpsh_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int bfpsh_set_physics(bfpsh_ctx* ctx, double nu, int forcing, double famplitude, double fk0, double fk1, int fmode) {
    if (!ctx) return BFPSH_EINVAL;
    if (forcing < 0 || forcing > 3) return fail(ctx, BFPSH_EINVAL, "bad forcing type");
    if (forcing == BFPSH_FORCING_FIXED_ENERGY && !ctx->d_famp) {
        HIPCHK(hipMalloc((void**)&ctx->d_famp, 3 * sizeof(double)));
        ctx->bytes += 3 * sizeof(double);
    }
    if (forcing == BFPSH_FORCING_KOLMOGOROV && (fmode < 0 || fmode >= ctx->n))
        return fail(ctx, BFPSH_EINVAL, "fmode out of range");
    ctx->sp.nu = nu;
    ctx->sp.forcing = forcing;
    ctx->sp.famp = famplitude;
    ctx->sp.fk0 = fk0;
    ctx->sp.fk1 = fk1;
    ctx->sp.fmode = fmode;
    return BFPSH_OK;
}

int bfpsh_kspace_info(bfpsh_ctx* ctx, double* out4, double* kshell, int64_t* nshell) {
    if (!ctx || !out4) return BFPSH_EINVAL;
    out4[0] = ctx->kM;
    out4[1] = ctx->kM2;
    out4[2] = ctx->dk;
    out4[3] = ctx->nshells;
    if (kshell) std::copy(ctx->kshell.begin(), ctx->kshell.end(), kshell);
    if (nshell) std::copy(ctx->nshell.begin(), ctx->nshell.end(), nshell);
    return BFPSH_OK;
}

int bfpsh_put_cfield(bfpsh_ctx* ctx, int which, const void* host) {
    CTX_CHECK();
    if (!host || (which != BFPSH_CVORTICITY && which != BFPSH_CVELOCITY))
        return fail(ctx, BFPSH_EINVAL, "only the vorticity or the velocity can be put");
    void* dst = ctx->v0;
    if (which == BFPSH_CVELOCITY) {
        // cvelocity lives in work planes 0..2 (the step reuses them)
        int r = ensure_work(ctx, 6);
        if (r) return r;
        dst = ctx->work;
    }
    const size_t bytes = 3 * ctx->ps * 2 * ctx->es;
    std::vector<unsigned char> tmp(bytes);
    if (ctx->prec == BFPSH_FP32) host_to_planar<float>((const float*)host, (float*)tmp.data(), ctx->n);
    else host_to_planar<double>((const double*)host, (double*)tmp.data(), ctx->n);
    HIPCHK(hipMemcpyAsync(dst, tmp.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (which == BFPSH_CVORTICITY) ctx->ru_valid = false;
    return BFPSH_OK;
}

int bfpsh_compute_vorticity(bfpsh_ctx* ctx) {
    CTX_CHECK();
    return PREC_CALL(compute_vorticity, ctx);
}

int bfpsh_impose_zero_modes(bfpsh_ctx* ctx) {
    CTX_CHECK();
    // the k = 0 mode is element 0 of each plane: cvelocity (work 0..2), v[0]
    // (cvorticity), v[1], v[2]
    const size_t cb = 2 * ctx->es, pb = ctx->ps * cb;
    void* bufs[4] = {ctx->work, ctx->v0, ctx->v1, ctx->v2};
    for (void* b : bufs) {
        if (!b) continue;
        for (int c = 0; c < 3; c++)
            HIPCHK(hipMemsetAsync(static_cast<unsigned char*>(b) + c * pb, 0, cb, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return BFPSH_OK;
}

int bfpsh_get_cfield(bfpsh_ctx* ctx, int which, void* host) {
    CTX_CHECK();
    if (!host) return fail(ctx, BFPSH_EINVAL, "null buffer");
    const void* src;
    if (which == BFPSH_CVORTICITY) {
        src = ctx->v0;
    } else if (which == BFPSH_CVELOCITY) {
        int r = PREC_CALL(spectral_velocity, ctx);
        if (r) return r;
        src = ctx->work;
    } else {
        return fail(ctx, BFPSH_EINVAL, "bad field selector");
    }
    const size_t bytes = 3 * ctx->ps * 2 * ctx->es;
    std::vector<unsigned char> tmp(bytes);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ctx->prec == BFPSH_FP32) planar_to_host<float>((const float*)tmp.data(), (float*)host, ctx->n);
    else planar_to_host<double>((const double*)tmp.data(), (double*)host, ctx->n);
    return BFPSH_OK;
}

int bfpsh_get_rfield(bfpsh_ctx* ctx, int which, void* host) {
    CTX_CHECK();
    if (!host) return fail(ctx, BFPSH_EINVAL, "null buffer");
    const void* src = which == BFPSH_RVELOCITY ? ctx->ru : (which == BFPSH_RGRADIENT ? ctx->rgrad : nullptr);
    const int nc = which == BFPSH_RVELOCITY ? 3 : 9;
    if (!src) return fail(ctx, BFPSH_ESTATE, "real field not computed");
    const size_t n = ctx->n, rps = 2 * ctx->ps, rys = 2 * row_stride(ctx->n);
    std::vector<unsigned char> tmp(nc * rps * ctx->es);
    HIPCHK(hipMemcpyAsync(tmp.data(), src, tmp.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // device [c][y][z][x] -> host [z][y][x][c]
    for (size_t y = 0; y < n; y++)
        for (size_t z = 0; z < n; z++)
            for (size_t x = 0; x < n; x++)
                for (int c = 0; c < nc; c++) {
                    const size_t di = c * rps + y * rys + z * n + x;
                    const size_t hi = ((z * n + y) * n + x) * nc + c;
                    if (ctx->prec == BFPSH_FP32) ((float*)host)[hi] = ((const float*)tmp.data())[di];
                    else ((double*)host)[hi] = ((const double*)tmp.data())[di];
                }
    return BFPSH_OK;
}

int bfpsh_load_projection(bfpsh_ctx* ctx) {
    CTX_CHECK();
    return PREC_CALL(load_projection, ctx);
}

int bfpsh_step(bfpsh_ctx* ctx, double dt) {
    CTX_CHECK();
    return PREC_CALL(step, ctx, dt, false);
}

int bfpsh_velocity_rspace(bfpsh_ctx* ctx) {
    CTX_CHECK();
    return PREC_CALL(velocity_rspace, ctx);
}

int bfpsh_gradient_rspace(bfpsh_ctx* ctx) {
    CTX_CHECK();
    return PREC_CALL(gradient_rspace, ctx);
}

int bfpsh_cospectrum(bfpsh_ctx* ctx, int a, int b, double* out) {
    CTX_CHECK();
    if (!out) return fail(ctx, BFPSH_EINVAL, "null output");
    // velocity needs its own buffer when both operands are used
    const bool need_u = a == BFPSH_CVELOCITY || b == BFPSH_CVELOCITY;
    if ((a != BFPSH_CVORTICITY && a != BFPSH_CVELOCITY) || (b != BFPSH_CVORTICITY && b != BFPSH_CVELOCITY))
        return fail(ctx, BFPSH_EINVAL, "bad field selector");
    if (need_u) {
        int r = PREC_CALL(spectral_velocity, ctx);
        if (r) return r;
    }
    const void* pa = a == BFPSH_CVORTICITY ? ctx->v0 : ctx->work;
    const void* pb = b == BFPSH_CVORTICITY ? ctx->v0 : ctx->work;
    return PREC_CALL(cospectrum, ctx, pa, pb, out);
}

static bool field_ok(bfpsh_ctx* ctx, int id) {
    return id >= 0 && id < int(ctx->fields.size()) && ctx->fields[id] != nullptr;
}

int bfpsh_field_create(bfpsh_ctx* ctx, int ncomp, int* id) {
    CTX_CHECK();
    if (!id || (ncomp != 1 && ncomp != 3 && ncomp != 9)) return fail(ctx, BFPSH_EINVAL, "ncomp must be 1, 3 or 9");
    void* p = nullptr;
    const size_t bytes = size_t(ncomp) * ctx->ps * 2 * ctx->es;
    HIPCHK(hipMalloc(&p, bytes));
    HIPCHK(hipMemsetAsync(p, 0, bytes, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->fields.push_back(p);
    ctx->field_nc.push_back(ncomp);
    ctx->field_real.push_back(0);
    ctx->bytes += bytes;
    *id = int(ctx->fields.size()) - 1;
    return BFPSH_OK;
}

int bfpsh_field_destroy(bfpsh_ctx* ctx, int id) {
    CTX_CHECK();
    if (!field_ok(ctx, id)) return fail(ctx, BFPSH_EINVAL, "bad field id");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipFree(ctx->fields[id]));
    ctx->bytes -= size_t(ctx->field_nc[id]) * ctx->ps * 2 * ctx->es;
    ctx->fields[id] = nullptr;
    return BFPSH_OK;
}

int bfpsh_field_ift(bfpsh_ctx* ctx, int id) {
    CTX_CHECK();
    if (!field_ok(ctx, id)) return fail(ctx, BFPSH_EINVAL, "bad field id");
    ctx->field_real[id] = 1;
    return PREC_CALL(field_ift, ctx, ctx->fields[id], ctx->field_nc[id]);
}

int bfpsh_field_dft(bfpsh_ctx* ctx, int id) {
    CTX_CHECK();
    if (!field_ok(ctx, id)) return fail(ctx, BFPSH_EINVAL, "bad field id");
    ctx->field_real[id] = 0;
    return PREC_CALL(field_dft, ctx, ctx->fields[id], ctx->field_nc[id]);
}

int bfpsh_field_symmetrize(bfpsh_ctx* ctx, int id) {
    CTX_CHECK();
    if (!field_ok(ctx, id)) return fail(ctx, BFPSH_EINVAL, "bad field id");
    const int n = ctx->n;
    if (ctx->prec == BFPSH_FP32)
        LAUNCH("symmetrize", (k_symmetrize<float><<<(n * n + 255) / 256, 256, 0, ctx->stream>>>((float2*)ctx->fields[id], n, ctx->field_nc[id], 0)));
    else
        LAUNCH("symmetrize", (k_symmetrize<double><<<(n * n + 255) / 256, 256, 0, ctx->stream>>>((double2*)ctx->fields[id], n, ctx->field_nc[id], 0)));
    return BFPSH_OK;
}

int bfpsh_field_compute_gradient(bfpsh_ctx* ctx, int src_id, int dst_id) {
    CTX_CHECK();
    if (!field_ok(ctx, src_id) || !field_ok(ctx, dst_id) || src_id == dst_id)
        return fail(ctx, BFPSH_EINVAL, "bad field ids");
    const int ncs = ctx->field_nc[src_id], ncd = ctx->field_nc[dst_id];
    if (!((ncs == 1 && ncd == 3) || (ncs == 3 && ncd == 9)))
        return fail(ctx, BFPSH_EINVAL, "gradient maps ONE to THREE or THREE to THREExTHREE");
    if (ctx->field_real[src_id]) return fail(ctx, BFPSH_ESTATE, "source field is in real space (call dft first)");
    const size_t cnt = size_t(ctx->n) * ctx->n * (ctx->n / 2);
    const unsigned grid = unsigned((cnt + 255) / 256);
    if (ctx->prec == BFPSH_FP32)
        LAUNCH("gradient", (k_gradient<float><<<grid, 256, 0, ctx->stream>>>((const float2*)ctx->fields[src_id],
                                                                             (float2*)ctx->fields[dst_id], geom(ctx), ncs)));
    else
        LAUNCH("gradient", (k_gradient<double><<<grid, 256, 0, ctx->stream>>>((const double2*)ctx->fields[src_id],
                                                                              (double2*)ctx->fields[dst_id], geom(ctx), ncs)));
    ctx->field_real[dst_id] = 0;
    return BFPSH_OK;
}

int bfpsh_field_put(bfpsh_ctx* ctx, int id, int real_space, const void* host) {
    CTX_CHECK();
    if (!field_ok(ctx, id) || !host) return fail(ctx, BFPSH_EINVAL, "bad field id or buffer");
    const int nc = ctx->field_nc[id];
    const size_t bytes = size_t(nc) * ctx->ps * 2 * ctx->es;
    std::vector<unsigned char> tmp(bytes, 0);
    if (real_space) {
        if (ctx->prec == BFPSH_FP32) host_real_to_planar<float>((const float*)host, (float*)tmp.data(), ctx->n, nc);
        else host_real_to_planar<double>((const double*)host, (double*)tmp.data(), ctx->n, nc);
    } else {
        if (ctx->prec == BFPSH_FP32) host_to_planar<float>((const float*)host, (float*)tmp.data(), ctx->n, nc);
        else host_to_planar<double>((const double*)host, (double*)tmp.data(), ctx->n, nc);
    }
    HIPCHK(hipMemcpyAsync(ctx->fields[id], tmp.data(), bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->field_real[id] = real_space ? 1 : 0;
    return BFPSH_OK;
}

int bfpsh_field_get(bfpsh_ctx* ctx, int id, int real_space, void* host) {
    CTX_CHECK();
    if (!field_ok(ctx, id) || !host) return fail(ctx, BFPSH_EINVAL, "bad field id or buffer");
    const int nc = ctx->field_nc[id];
    const size_t bytes = size_t(nc) * ctx->ps * 2 * ctx->es;
    std::vector<unsigned char> tmp(bytes);
    HIPCHK(hipMemcpyAsync(tmp.data(), ctx->fields[id], bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (real_space) {
        if (ctx->prec == BFPSH_FP32) planar_to_host_real<float>((const float*)tmp.data(), (float*)host, ctx->n, nc);
        else planar_to_host_real<double>((const double*)tmp.data(), (double*)host, ctx->n, nc);
    } else {
        if (ctx->prec == BFPSH_FP32) planar_to_host<float>((const float*)tmp.data(), (float*)host, ctx->n, nc);
        else planar_to_host<double>((const double*)tmp.data(), (double*)host, ctx->n, nc);
    }
    return BFPSH_OK;
}

int bfpsh_field_copy_from(bfpsh_ctx* ctx, int id, int which) {
    CTX_CHECK();
    if (!field_ok(ctx, id) || ctx->field_nc[id] != 3) return fail(ctx, BFPSH_EINVAL, "needs a 3-component field");
    const void* src;
    if (which == BFPSH_CVORTICITY) {
        src = ctx->v0;
    } else if (which == BFPSH_CVELOCITY) {
        int r = PREC_CALL(spectral_velocity, ctx);
        if (r) return r;
        src = ctx->work;
    } else {
        return fail(ctx, BFPSH_EINVAL, "bad field selector");
    }
    HIPCHK(hipMemcpyAsync(ctx->fields[id], src, 3 * ctx->ps * 2 * ctx->es, hipMemcpyDeviceToDevice, ctx->stream));
    ctx->field_real[id] = 0;
    return BFPSH_OK;
}

int bfpsh_field_cospectrum(bfpsh_ctx* ctx, int id_a, int id_b, double* out) {
    CTX_CHECK();
    if (!out || !field_ok(ctx, id_a) || !field_ok(ctx, id_b) || ctx->field_nc[id_a] != 3 || ctx->field_nc[id_b] != 3)
        return fail(ctx, BFPSH_EINVAL, "needs two 3-component fields");
    return PREC_CALL(cospectrum, ctx, ctx->fields[id_a], ctx->fields[id_b], out);
}

// The same shell sums with no allocation: partial sums in this device's
// copy of `scratch` (cleared by a fill kernel first), copied to `out`.
int bfpsh_field_cospectrum_noalloc(bfpsh_ctx* ctx, int id_a, int id_b, double* out) {
    CTX_CHECK();
    if (!out || !field_ok(ctx, id_a) || !field_ok(ctx, id_b) || ctx->field_nc[id_a] != 3 || ctx->field_nc[id_b] != 3)
        return fail(ctx, BFPSH_EINVAL, "needs two 3-component fields");
    if (ctx->nshells > kScratchMaxShells)
        return fail(ctx, BFPSH_EINVAL, "cospectrum of " + std::to_string(ctx->nshells) + " shells exceeds the " +
                                           std::to_string(kScratchMaxShells) + "-shell fixed scratch");
    if (ctx->device < 0 || ctx->device >= kScratchMaxDevices || !scratch_dev[ctx->device])
        return fail(ctx, BFPSH_ESTATE, "cospectrum scratch not resolved on this device");
    double* d = scratch_dev[ctx->device];
    const int nv = ctx->nshells * 9;
    const size_t lds = size_t(nv) * sizeof(double);
    if (lds > 64 * 1024) return fail(ctx, BFPSH_EINVAL, "too many shells for the cospectrum kernel");
    k_fill_zero_d<<<(nv + 255) / 256, 256, 0, ctx->stream>>>(d, nv);
    HIPCHK(hipGetLastError());
    if (ctx->prec == BFPSH_FP32)
        k_cospectrum<float><<<1024, 256, lds, ctx->stream>>>((const float2*)ctx->fields[id_a],
                                                             (const float2*)ctx->fields[id_b], d, geom(ctx), ctx->dk,
                                                             ctx->nshells);
    else
        k_cospectrum<double><<<1024, 256, lds, ctx->stream>>>((const double2*)ctx->fields[id_a],
                                                              (const double2*)ctx->fields[id_b], d, geom(ctx), ctx->dk,
                                                              ctx->nshells);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, d, lds, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return BFPSH_OK;
}

int bfpsh_field_normalize(bfpsh_ctx* ctx, int id) {
    CTX_CHECK();
    if (!field_ok(ctx, id)) return fail(ctx, BFPSH_EINVAL, "bad field id");
    return PREC_CALL(normalize, ctx, ctx->fields[id], ctx->field_nc[id]);
}

int bfpsh_field_rspace_stats(bfpsh_ctx* ctx, int id, int nbins, const double* max_estimate, double* moments,
                             int64_t* hist) {
    CTX_CHECK();
    if (!field_ok(ctx, id)) return fail(ctx, BFPSH_EINVAL, "bad field id");
    if (!max_estimate || !moments || !hist) return fail(ctx, BFPSH_EINVAL, "null argument");
    if (nbins < 1 || nbins > 4096) return fail(ctx, BFPSH_EINVAL, "nbins must be in [1, 4096]");
    if (!ctx->field_real[id]) return fail(ctx, BFPSH_ESTATE, "field is not in real space (call ift first)");
    for (int i = 0; i < (ctx->field_nc[id] == 3 ? 4 : ctx->field_nc[id]); i++)
        if (!(max_estimate[i] > 0)) return fail(ctx, BFPSH_EINVAL, "max_estimate values must be positive");
    return PREC_CALL(rspace_stats, ctx, ctx->fields[id], ctx->field_nc[id], nbins, max_estimate, moments, hist);
}

int bfpsh_compute_pressure(bfpsh_ctx* ctx, int id) {
    CTX_CHECK();
    if (!field_ok(ctx, id) || ctx->field_nc[id] != 1) return fail(ctx, BFPSH_EINVAL, "needs a 1-component field");
    ctx->field_real[id] = 0;
    return PREC_CALL(pressure, ctx, ctx->fields[id]);
}

int bfpsh_compute_lagrangian_acceleration(bfpsh_ctx* ctx, int id) {
    CTX_CHECK();
    if (!field_ok(ctx, id) || ctx->field_nc[id] != 3) return fail(ctx, BFPSH_EINVAL, "needs a 3-component field");
    ctx->field_real[id] = 0;
    return PREC_CALL(lagrangian_acceleration, ctx, ctx->fields[id]);
}

// Slots in global-index order, on the host (order) and as a rank per slot
// on the device.  The storage order of tracers never changes (sorting for
// locality only permutes an index array), so this is set once per put.
static int set_order(bfpsh_ctx* ctx) {
    ctx->order.resize(ctx->np);
    std::iota(ctx->order.begin(), ctx->order.end(), int64_t(0));
    std::stable_sort(ctx->order.begin(), ctx->order.end(),
                     [&](int64_t a, int64_t b) { return ctx->idx[a] < ctx->idx[b]; });
    if (ctx->np == 0) return BFPSH_OK;
    std::vector<uint32_t> rank(ctx->np);
    for (int64_t i = 0; i < ctx->np; i++) rank[ctx->order[i]] = uint32_t(i);
    if (!ctx->d_rank) {
        HIPCHK(hipMalloc((void**)&ctx->d_rank, ctx->np * sizeof(uint32_t)));
        ctx->bytes += ctx->np * sizeof(uint32_t);
    }
    HIPCHK(hipMemcpyAsync(ctx->d_rank, rank.data(), ctx->np * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return BFPSH_OK;
}

int bfpsh_particles_create(bfpsh_ctx* ctx, int64_t np, int n, int m, int nsteps, int step_idx) {
    CTX_CHECK();
    if (np < 0 || n < 1 || n > BFPS_INTERP_NMAX || m < 0 || m > BFPS_INTERP_MMAX || nsteps < 1 || nsteps > 6)
        return fail(ctx, BFPSH_EINVAL, "bad tracer parameters (n in [1,10], m in [0,2], nsteps in [1,6])");
    if (2 * n + 2 > ctx->n) return fail(ctx, BFPSH_EINVAL, "interpolation stencil wider than the grid");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (void* p : {(void*)ctx->d_pos, (void*)ctx->d_rhs, (void*)ctx->d_tab, (void*)ctx->d_rank, (void*)ctx->d_samp})
        if (p) HIPCHK(hipFree(p));
    ctx->d_rank = nullptr;
    ctx->d_samp = nullptr;
    {
        int r = bfpsh_part_release(ctx);
        if (r) return r;
    }
    ctx->d_pos = ctx->d_rhs = ctx->d_tab = nullptr;
    ctx->np = np;
    ctx->pn = n;
    ctx->pm = m;
    {
        // kernel choice for the velocity loop: automatic, or forced to one
        // form (the tests run each form on the same case)
        const char* tk = std::getenv("BFPSH_TRACER_KERNEL");
        ctx->tracer_kernel = !tk ? 0
                             : !std::strcmp(tk, "brick")  ? 1
                             : !std::strcmp(tk, "gather") ? 2
                             : !std::strcmp(tk, "column") ? 3
                                                          : 0;
        ctx->tracer_form = 0;
    }
    ctx->nsteps = nsteps;
    ctx->step_idx = step_idx;
    ctx->head = 0;
    ctx->head_zero = false;
    ctx->add_head = false;
    const int k = (n - 1) * (BFPS_INTERP_MMAX + 1) + m;
    ctx->deg = bfps_interp_degree[k];
    const size_t tabn = size_t(3) * (2 * n + 2) * (ctx->deg + 1);
    HIPCHK(hipMalloc((void**)&ctx->d_tab, tabn * sizeof(double)));
    HIPCHK(hipMemcpyAsync(ctx->d_tab, bfps_interp_coef + bfps_interp_offset[k], tabn * sizeof(double),
                          hipMemcpyHostToDevice, ctx->stream));
    if (np > 0) {
        HIPCHK(hipMalloc((void**)&ctx->d_pos, np * 3 * sizeof(double)));
        HIPCHK(hipMalloc((void**)&ctx->d_rhs, size_t(nsteps) * np * 3 * sizeof(double)));
        HIPCHK(hipMemsetAsync(ctx->d_pos, 0, np * 3 * sizeof(double), ctx->stream));
        HIPCHK(hipMemsetAsync(ctx->d_rhs, 0, size_t(nsteps) * np * 3 * sizeof(double), ctx->stream));
    }
    ctx->idx.resize(np);
    std::iota(ctx->idx.begin(), ctx->idx.end(), int64_t(0));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return set_order(ctx);
}

int bfpsh_particles_put(bfpsh_ctx* ctx, const double* pos, const double* rhs, const int64_t* idx) {
    CTX_CHECK();
    if (!pos) return fail(ctx, BFPSH_EINVAL, "null positions");
    if (ctx->np == 0) return BFPSH_OK;
    const size_t pb = ctx->np * 3 * sizeof(double);
    HIPCHK(hipMemcpyAsync(ctx->d_pos, pos, pb, hipMemcpyHostToDevice, ctx->stream));
    ctx->head = 0;
    ctx->add_head = false;
    if (rhs) {
        HIPCHK(hipMemcpyAsync(ctx->d_rhs, rhs, pb * ctx->nsteps, hipMemcpyHostToDevice, ctx->stream));
        for (int64_t i = 0; i < ctx->np * 3; i++)
            if (rhs[i] != 0.0) { ctx->add_head = true; break; }
    } else {
        HIPCHK(hipMemsetAsync(ctx->d_rhs, 0, pb * ctx->nsteps, ctx->stream));
    }
    ctx->head_zero = !ctx->add_head;
    ctx->sort_age = 1 << 30;  // positions replaced: re-sort before the next loop
    if (idx) std::copy(idx, idx + ctx->np, ctx->idx.begin());
    else std::iota(ctx->idx.begin(), ctx->idx.end(), int64_t(0));
    return set_order(ctx);
}

int bfpsh_particles_get(bfpsh_ctx* ctx, double* pos, double* rhs, int64_t* idx, int* step_idx) {
    CTX_CHECK();
    if (step_idx) *step_idx = ctx->step_idx;
    if (ctx->np == 0) return BFPSH_OK;
    const int64_t np = ctx->np;
    const std::vector<int64_t>& order = ctx->order;
    if (int64_t(order.size()) != np) return fail(ctx, BFPSH_ESTATE, "tracer state not set (particles_put)");
    std::vector<double> tmp(np * 3);
    if (pos) {
        HIPCHK(hipMemcpyAsync(tmp.data(), ctx->d_pos, np * 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        for (int64_t i = 0; i < np; i++)
            for (int d = 0; d < 3; d++) pos[i * 3 + d] = tmp[order[i] * 3 + d];
    }
    if (rhs) {
        for (int h = 0; h < ctx->nsteps; h++) {
            const int slot = (ctx->head + h) % ctx->nsteps;
            if (h == 0 && ctx->head_zero) {
                std::fill(rhs, rhs + np * 3, 0.0);
                continue;
            }
            HIPCHK(hipMemcpyAsync(tmp.data(), ctx->d_rhs + size_t(slot) * np * 3, np * 3 * sizeof(double),
                                  hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            for (int64_t i = 0; i < np; i++)
                for (int d = 0; d < 3; d++) rhs[(size_t(h) * np + i) * 3 + d] = tmp[order[i] * 3 + d];
        }
    }
    if (idx)
        for (int64_t i = 0; i < np; i++) idx[i] = ctx->idx[order[i]];
    return BFPSH_OK;
}

int bfpsh_particles_complete_loop(bfpsh_ctx* ctx, double dt) {
    CTX_CHECK();
    if (ctx->nsteps == 0) return fail(ctx, BFPSH_ESTATE, "no tracers");
    if (!ctx->ru_valid) {
        int r = PREC_CALL(velocity_rspace, ctx);
        if (r) return r;
    }
    return bfpsh_part_loop(ctx, dt);
}

// out[rank[p]][c] = in[p][c]: device rows into global-index order
static __global__ void k_rank_rows(const double* __restrict__ in, double* __restrict__ out,
                                   const uint32_t* __restrict__ rank, int64_t np, int nc) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= np * nc) return;
    const int64_t p = i / nc;
    out[int64_t(rank[p]) * nc + (i - p * nc)] = in[i];
}

// sample `nc` real planes [c][y][z][x] at every tracer; out [np][nc] in
// global-index order
static int sample_to_host(bfpsh_ctx* ctx, const void* f, int nc, double* out) {
    if (ctx->np == 0) return BFPSH_OK;
    if (int64_t(ctx->order.size()) != ctx->np) return fail(ctx, BFPSH_ESTATE, "tracer state not set (particles_put)");
    if (!ctx->d_samp) {
        const size_t bytes = size_t(ctx->np) * 9 * 2 * sizeof(double);
        HIPCHK(hipMalloc((void**)&ctx->d_samp, bytes));
        ctx->bytes += bytes;
    }
    double* raw = ctx->d_samp;
    double* ranked = ctx->d_samp + size_t(ctx->np) * 9;
    int r = bfpsh_part_sample(ctx, f, nc, raw);
    if (r) return r;
    const int64_t tot = ctx->np * nc;
    LAUNCH("rank_rows", (k_rank_rows<<<unsigned((tot + 255) / 256), 256, 0, ctx->stream>>>(raw, ranked, ctx->d_rank,
                                                                                          ctx->np, nc)));
    HIPCHK(hipMemcpyAsync(out, ranked, size_t(tot) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return BFPSH_OK;
}

int bfpsh_particles_sample(bfpsh_ctx* ctx, int which, double* out) {
    CTX_CHECK();
    if (!out) return fail(ctx, BFPSH_EINVAL, "null output");
    if (which != BFPSH_RVELOCITY && which != BFPSH_RGRADIENT) return fail(ctx, BFPSH_EINVAL, "bad field selector");
    if ((which == BFPSH_RVELOCITY && !ctx->ru) || (which == BFPSH_RGRADIENT && !ctx->rgrad))
        return fail(ctx, BFPSH_ESTATE, "real field not computed");
    if (ctx->nsteps == 0) return fail(ctx, BFPSH_ESTATE, "no tracers");
    return sample_to_host(ctx, which == BFPSH_RVELOCITY ? ctx->ru : ctx->rgrad, which == BFPSH_RVELOCITY ? 3 : 9,
                          out);
}

int bfpsh_particles_sample_field(bfpsh_ctx* ctx, int id, double* out) {
    CTX_CHECK();
    if (!field_ok(ctx, id) || !out) return fail(ctx, BFPSH_EINVAL, "bad field id or buffer");
    if (ctx->nsteps == 0) return fail(ctx, BFPSH_ESTATE, "no tracers");
    // sample_compute_field of a one-component float field does nothing
    // (bfps/cpp/particles/particles_system.hpp:147-150)
    if (ctx->field_nc[id] == 1 && ctx->prec == BFPSH_FP32) return BFPSH_OK;
    if (!ctx->field_real[id]) return fail(ctx, BFPSH_ESTATE, "field is not in real space (call ift first)");
    return sample_to_host(ctx, ctx->fields[id], ctx->field_nc[id], out);
}

int bfpsh_nsve_particles_step(bfpsh_ctx* ctx, double dt) {
    CTX_CHECK();
    return PREC_CALL(step, ctx, dt, ctx->nsteps > 0 && ctx->np > 0);
}

int bfpsh_synchronize(bfpsh_ctx* ctx) {
    CTX_CHECK();
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return BFPSH_OK;
}

int bfpsh_get_iteration(bfpsh_ctx* ctx, int* it) {
    if (!ctx || !it) return BFPSH_EINVAL;
    *it = ctx->iteration;
    return BFPSH_OK;
}
int bfpsh_set_iteration(bfpsh_ctx* ctx, int it) {
    if (!ctx) return BFPSH_EINVAL;
    ctx->iteration = it;
    return BFPSH_OK;
}

int bfpsh_time_steps(bfpsh_ctx* ctx, int nsteps, double dt, int with_particles, double* ms) {
    CTX_CHECK();
    if (!ms || nsteps < 0) return fail(ctx, BFPSH_EINVAL, "bad arguments");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, ctx->stream));
    const bool part = with_particles && ctx->nsteps > 0 && ctx->np > 0;
    for (int i = 0; i < nsteps; i++) {
        int r = PREC_CALL(step, ctx, dt, part);
        if (r) return r;
    }
    HIPCHK(hipEventRecord(b, ctx->stream));
    HIPCHK(hipEventSynchronize(b));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, a, b));
    *ms = t;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return BFPSH_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ pass wisdom
namespace {
std::mutex wisdom_mu;
std::map<std::string, PassWisdom> wisdom;
std::map<int, std::string> arch_names;  // device -> architecture (gcnArchName up to ':')
const char* const kPassKindName[BFPSH_PASS_KINDS] = {"z_inv", "y_inv", "z_fwd", "y_fwd", "bs", "y3"};
}  // namespace

std::string pass_wisdom_key(const bfpsh_ctx* ctx, int kind) {
    std::string arch;
    {
        std::lock_guard<std::mutex> lk(wisdom_mu);
        auto it = arch_names.find(ctx->device);
        if (it != arch_names.end()) arch = it->second;
    }
    if (arch.empty()) {
        hipDeviceProp_t prop;
        arch = "unknown";
        if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) {
            arch = prop.gcnArchName;
            arch = arch.substr(0, arch.find(':'));
        }
        std::lock_guard<std::mutex> lk(wisdom_mu);
        arch_names[ctx->device] = arch;
    }
    return arch + "/" + (ctx->prec == BFPSH_FP32 ? "f32" : "f64") + "/" + std::to_string(ctx->n) + "/" +
           kPassKindName[(kind >= 0 && kind < BFPSH_PASS_KINDS) ? kind : 0];
}

bool pass_wisdom_lookup(const std::string& key, PassWisdom* w) {
    std::lock_guard<std::mutex> lk(wisdom_mu);
    auto it = wisdom.find(key);
    if (it == wisdom.end()) return false;
    *w = it->second;
    return true;
}

void pass_wisdom_store(const std::string& key, const PassWisdom& w) {
    std::lock_guard<std::mutex> lk(wisdom_mu);
    wisdom[key] = w;
}

extern "C" {

int bfpsh_set_plan_rigor(bfpsh_ctx* ctx, int rigor) {
    CTX_CHECK();
    if (rigor < BFPSH_ESTIMATE || rigor > BFPSH_PATIENT) return fail(ctx, BFPSH_EINVAL, "bad plan rigor");
    return PREC_CALL(tune_passes, ctx, rigor);
}

int bfpsh_get_pass_shape(bfpsh_ctx* ctx, int kind, int* e, int* b, int* persistent, int* index, int* nshapes) {
    CTX_CHECK();
    if (kind < 0 || kind >= BFPSH_PASS_KINDS) return fail(ctx, BFPSH_EINVAL, "bad pass kind");
    const PassShapes sh = ops_for(ctx)->pass_shapes(kind);
    const int v = ctx->pass_var[kind] >= 0 && ctx->pass_var[kind] < sh.n ? ctx->pass_var[kind] : 0;
    if (e) *e = sh.v[v].e;
    if (b) *b = sh.v[v].b;
    if (persistent) *persistent = sh.v[v].pp;
    if (index) *index = v;
    if (nshapes) *nshapes = sh.n;
    return BFPSH_OK;
}

int bfpsh_set_pass_shape(bfpsh_ctx* ctx, int kind, int index) {
    CTX_CHECK();
    if (kind < 0 || kind >= BFPSH_PASS_KINDS) return fail(ctx, BFPSH_EINVAL, "bad pass kind");
    const PassShapes sh = ops_for(ctx)->pass_shapes(kind);
    if (index < 0 || index >= sh.n) return fail(ctx, BFPSH_EINVAL, "pass shape index out of range");
    ctx->pass_var[kind] = index;
    return BFPSH_OK;
}

int bfpsh_wisdom_import(const char* path) {
    if (!path) return BFPSH_EINVAL;
    FILE* f = std::fopen(path, "r");
    if (!f) return BFPSH_OK;  // no wisdom yet
    std::map<std::string, PassWisdom> got;
    char line[512];
    int rc = BFPSH_OK;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        char key[256];
        PassWisdom w;
        if (std::sscanf(line, "%255s %d %d %d %lf %d", key, &w.e, &w.b, &w.pp, &w.ms, &w.rigor) != 6 || w.e <= 0 ||
            w.b <= 0 || (w.pp != 0 && w.pp != 1) || w.rigor < BFPSH_ESTIMATE || w.rigor > BFPSH_PATIENT) {
            rc = BFPSH_EINVAL;
            break;
        }
        got[key] = w;
    }
    std::fclose(f);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(wisdom_mu);
    for (auto& kv : got) wisdom[kv.first] = kv.second;
    return BFPSH_OK;
}

int bfpsh_wisdom_export(const char* path) {
    if (!path) return BFPSH_EINVAL;
    std::map<std::string, PassWisdom> snap;
    {
        std::lock_guard<std::mutex> lk(wisdom_mu);
        snap = wisdom;
    }
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f) return BFPSH_EINVAL;
    std::fprintf(f, "# bfps_amd pass wisdom: <arch>/<precision>/<N>/<pass kind> <values per thread> <columns> "
                    "<persistent> <ms> <rigor>\n");
    for (auto& kv : snap)
        std::fprintf(f, "%s %d %d %d %.6f %d\n", kv.first.c_str(), kv.second.e, kv.second.b, kv.second.pp, kv.second.ms,
                     kv.second.rigor);
    if (std::fclose(f) != 0 || std::rename(tmp.c_str(), path) != 0) return BFPSH_EINVAL;
    return BFPSH_OK;
}

int bfpsh_wisdom_forget(void) {
    std::lock_guard<std::mutex> lk(wisdom_mu);
    wisdom.clear();
    return BFPSH_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ ensembles
// Independent boxes advanced together: boxes of one size form a class whose
// 18-plane buffers (v0 | v1 | v2 | work | ru) sit in one slab, so each field
// kernel of a step covers the whole class in one launch (an extra grid
// dimension over the boxes, per-box physics from a device table).  The field
// launches of all classes are captured once as a HIP graph and replayed each
// step; the tracers of each box follow with the stage-1 velocity.
struct bfpsh_ensemble {
    int device = 0, prec = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<bfpsh_ctx*> boxes;
    struct Cls {
        int n = 0;
        std::vector<int> members;
        void* slab = nullptr;
        size_t bstride = 0;            // complex elements between boxes
        StepParams* d_sp = nullptr;    // per-box parameters
        double* d_famp = nullptr;      // 3 band amplitudes per box
        std::vector<StepParams> sp;    // host copy of d_sp
    };
    std::vector<Cls> cls;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool captured_ru = false;          // the graph writes the real velocity
    std::vector<char> captured_fixed;  // per class: the graph holds the band-amplitude launches
    static constexpr int NPB = 4;      // neighbour counts batched in one step
    PartBatch pb[NPB];                 // tracer-loop groups, one per neighbour count
    int ngroups = 0;                   // groups the captured graph launches
    std::vector<uint64_t> captured_layout;  // their launch shapes and tables
};

static int efail(bfpsh_ensemble* e, int code, const std::string& msg) {
    e->err = msg;
    return code;
}
#define EHIPCHK(x)                                                              \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) return efail(e, BFPSH_EHIP, hipGetErrorString(e_)); \
    } while (0)

static void ensemble_free(bfpsh_ensemble* e) {
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (PartBatch& pb : e->pb) bfpsh_part_batch_free(pb);
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    for (bfpsh_ctx* c : e->boxes)
        if (c) bfpsh_destroy(c);
    for (auto& c : e->cls) {
        for (void* p : {c.slab, (void*)c.d_sp, (void*)c.d_famp})
            if (p) (void)hipFree(p);
    }
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

// Upload the per-box parameters (physics of each box context, its dt);
// returns whether anything changed.
static int ensemble_params(bfpsh_ensemble* e, const double* dt, bool* changed) {
    *changed = false;
    for (auto& c : e->cls) {
        bool diff = false;
        for (size_t j = 0; j < c.members.size(); j++) {
            StepParams p = e->boxes[c.members[j]]->sp;
            p.dt = dt[c.members[j]];
            p.famp_dev = c.d_famp + 3 * j;
            if (std::memcmp(&p, &c.sp[j], sizeof(StepParams)) != 0) {
                c.sp[j] = p;
                diff = true;
            }
        }
        if (diff) {
            EHIPCHK(hipStreamSynchronize(e->stream));
            EHIPCHK(hipMemcpy(c.d_sp, c.sp.data(), c.sp.size() * sizeof(StepParams), hipMemcpyHostToDevice));
            *changed = true;
        }
    }
    return BFPSH_OK;
}

// some box of the class uses the fixed energy-injection forcing, whose band
// amplitudes the graph must then compute each stage
static bool class_any_fixed(const bfpsh_ensemble::Cls& c) {
    for (const StepParams& p : c.sp)
        if (p.forcing == BFPSH_FORCING_FIXED_ENERGY) return true;
    return false;
}

// the captured graph no longer matches the boxes' forcing choices
static bool fixed_changed(const bfpsh_ensemble* e) {
    if (e->captured_fixed.size() != e->cls.size()) return true;
    for (size_t ci = 0; ci < e->cls.size(); ci++)
        if (bool(e->captured_fixed[ci]) != class_any_fixed(e->cls[ci])) return true;
    return false;
}

// what the captured tracer launches depend on: per group its neighbour
// count, precision, grid, LDS, table and state addresses
static std::vector<uint64_t> tracer_layout(const bfpsh_ensemble* e, int ngroups) {
    std::vector<uint64_t> v;
    for (int g = 0; g < ngroups; g++) {
        const PartBatch& b = e->pb[g];
        v.insert(v.end(), {uint64_t(b.nbr), uint64_t(b.prec), uint64_t(b.blocks), uint64_t(b.nboxes), uint64_t(b.lds),
                           uint64_t(reinterpret_cast<uintptr_t>(b.d)), uint64_t(reinterpret_cast<uintptr_t>(b.d_dyn))});
    }
    return v;
}

// One graph per step: the field kernels of every class (the stage-1 real
// velocity of each box kept in its ru when with_ru) and, with the tracers,
// the loop of each neighbour-count group and the advance of its device
// Adams-Bashforth state.
static int ensemble_capture(bfpsh_ensemble* e, bool with_ru, int ngroups) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
    // no per-launch profiling events inside the graph
    std::vector<bool> prof(e->boxes.size());
    for (size_t i = 0; i < e->boxes.size(); i++) {
        prof[i] = e->boxes[i]->prof;
        e->boxes[i]->prof = false;
    }
    EHIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = BFPSH_OK;
    e->captured_fixed.assign(e->cls.size(), 0);
    for (size_t ci = 0; ci < e->cls.size(); ci++) {
        auto& c = e->cls[ci];
        bfpsh_ctx* b0 = e->boxes[c.members[0]];
        BatchDesc d;
        d.nbox = int(c.members.size());
        d.bstride = c.bstride;
        d.spb = c.d_sp;
        d.any_fixed = class_any_fixed(c);
        e->captured_fixed[ci] = d.any_fixed;
        d.v0 = b0->v0;
        d.v1 = b0->v1;
        d.v2 = b0->v2;
        d.w = b0->work;
        d.ru = with_ru ? b0->ru : nullptr;
        rc = ops_for(b0)->batch_step(b0, d);
        if (rc) {
            e->err = b0->err;
            break;
        }
    }
    if (rc == BFPSH_OK && with_ru && ngroups > 0) {
        rc = bfpsh_part_ens_launch(e->pb, ngroups);
        if (rc) e->err = e->pb[0].ctx ? e->pb[0].ctx->err : "tracer launch failed";
    }
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    for (size_t i = 0; i < e->boxes.size(); i++) e->boxes[i]->prof = prof[i];
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ce != hipSuccess) return efail(e, BFPSH_EHIP, hipGetErrorString(ce));
    e->graph = g;
    EHIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    e->captured_ru = with_ru;
    e->ngroups = with_ru ? ngroups : 0;
    e->captured_layout = tracer_layout(e, e->ngroups);
    return BFPSH_OK;
}

extern "C" {

int bfpsh_ensemble_create(bfpsh_ensemble** out, int nboxes, const int* sizes, int precision, int dealias,
                          int device) {
    if (!out) return BFPSH_EINVAL;
    bfpsh_ensemble* e = new bfpsh_ensemble();
    *out = e;
    if (nboxes < 1 || !sizes) return efail(e, BFPSH_EINVAL, "need at least one box");
    if (precision != BFPSH_FP32 && precision != BFPSH_FP64) return efail(e, BFPSH_EINVAL, "bad precision");
    for (int i = 0; i < nboxes; i++)
        if (!supported_size(sizes[i])) return efail(e, BFPSH_EINVAL, "unsupported box size");
    e->device = device;
    e->prec = precision;
    EHIPCHK(hipSetDevice(device));
    EHIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    const size_t es = precision == BFPSH_FP32 ? 4 : 8;
    e->boxes.assign(nboxes, nullptr);
    std::vector<int> order(nboxes);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sizes[a] < sizes[b]; });
    for (int i : order) {
        if (e->cls.empty() || e->cls.back().n != sizes[i]) {
            e->cls.emplace_back();
            e->cls.back().n = sizes[i];
        }
        e->cls.back().members.push_back(i);
    }
    for (auto& c : e->cls) {
        const size_t ps = plane_size(c.n);
        const size_t box_bytes = 18 * ps * 2 * es;
        const size_t nb = c.members.size();
        c.bstride = 18 * ps;
        EHIPCHK(hipMalloc(&c.slab, nb * box_bytes));
        EHIPCHK(hipMemsetAsync(c.slab, 0, nb * box_bytes, e->stream));
        EHIPCHK(hipMalloc((void**)&c.d_sp, nb * sizeof(StepParams)));
        EHIPCHK(hipMalloc((void**)&c.d_famp, 3 * nb * sizeof(double)));
        EHIPCHK(hipMemsetAsync(c.d_famp, 0, 3 * nb * sizeof(double), e->stream));
        c.sp.assign(nb, StepParams{});
        for (size_t j = 0; j < nb; j++) {
            bfpsh_ctx* ctx = nullptr;
            int r = create_ctx(&ctx, c.n, c.n, c.n, 1.0, 1.0, 1.0, precision, dealias, device, e->stream,
                               (char*)c.slab + j * box_bytes);
            e->boxes[c.members[j]] = ctx;
            if (r) return efail(e, r, ctx ? ctx->err : "box creation failed");
            ctx->d_famp = c.d_famp + 3 * j;  // shared with the batch table (not freed by the box)
            ctx->force_coop = true;
            ctx->sort_every = 16;
            // the first step's upload always differs from this placeholder
            c.sp[j].nu = -1.0;
        }
    }
    EHIPCHK(hipStreamSynchronize(e->stream));
    return BFPSH_OK;
}

int bfpsh_ensemble_destroy(bfpsh_ensemble* e) {
    if (!e) return BFPSH_OK;
    (void)hipSetDevice(e->device);
    for (auto& c : e->cls)
        for (size_t j = 0; j < c.members.size(); j++) {
            bfpsh_ctx* ctx = e->boxes[c.members[j]];
            if (ctx) ctx->d_famp = nullptr;  // owned by the class table
        }
    ensemble_free(e);
    return BFPSH_OK;
}

const char* bfpsh_ensemble_last_error(const bfpsh_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

int bfpsh_ensemble_box(bfpsh_ensemble* e, int i, bfpsh_ctx** box) {
    if (!e || !box) return BFPSH_EINVAL;
    if (i < 0 || i >= int(e->boxes.size())) return efail(e, BFPSH_EINVAL, "box index out of range");
    *box = e->boxes[i];
    return BFPSH_OK;
}

int bfpsh_ensemble_step(bfpsh_ensemble* e, const double* dt, int with_particles) {
    if (!e || !dt) return BFPSH_EINVAL;
    EHIPCHK(hipSetDevice(e->device));
    bool changed = false;
    int r = ensemble_params(e, dt, &changed);
    if (r) return r;
    std::vector<bfpsh_ctx*> pbox;
    std::vector<double> pdt;
    if (with_particles)
        for (size_t i = 0; i < e->boxes.size(); i++) {
            bfpsh_ctx* c = e->boxes[i];
            if (c->nsteps > 0 && c->np > 0) {
                pbox.push_back(c);
                pdt.push_back(dt[i]);
            }
        }
    const bool any_part = !pbox.empty();
    int ngroups = 0;
    if (any_part) {
        // tracer tables and device AB state (rewritten only on a change),
        // and the periodic sorts of the boxes that are due
        r = bfpsh_part_ens_prepare(pbox, pdt, e->pb, bfpsh_ensemble::NPB, &ngroups);
        if (r) return efail(e, r, pbox[0]->err);
    }
    // parameter values reach the graph through the device tables; a box
    // switching to or from the fixed-injection forcing, or a change of the
    // tracer launch shapes, changes its launches
    if (!e->exec || e->captured_ru != any_part || (changed && fixed_changed(e)) ||
        tracer_layout(e, ngroups) != e->captured_layout) {
        r = ensemble_capture(e, any_part, ngroups);
        if (r) return r;
    }
    EHIPCHK(hipGraphLaunch(e->exec, e->stream));
    if (any_part) bfpsh_part_ens_advanced(pbox, e->pb, ngroups);
    for (bfpsh_ctx* c : e->boxes) {
        c->ru_valid = false;
        c->iteration++;
    }
    return BFPSH_OK;
}

int bfpsh_ensemble_step_separate(bfpsh_ensemble* e, const double* dt, int with_particles) {
    if (!e || !dt) return BFPSH_EINVAL;
    EHIPCHK(hipSetDevice(e->device));
    for (size_t i = 0; i < e->boxes.size(); i++) {
        bfpsh_ctx* ctx = e->boxes[i];
        ctx->sp.dt = dt[i];
        int r = PREC_CALL(step, ctx, dt[i], with_particles && ctx->nsteps > 0 && ctx->np > 0);
        if (r) return efail(e, r, ctx->err);
    }
    return BFPSH_OK;
}

int bfpsh_ensemble_time_steps(bfpsh_ensemble* e, int nsteps, const double* dt, int with_particles, double* ms) {
    if (!e || !ms || nsteps < 0) return BFPSH_EINVAL;
    EHIPCHK(hipSetDevice(e->device));
    hipEvent_t a, b;
    EHIPCHK(hipEventCreate(&a));
    EHIPCHK(hipEventCreate(&b));
    EHIPCHK(hipEventRecord(a, e->stream));
    for (int i = 0; i < nsteps; i++) {
        int r = bfpsh_ensemble_step(e, dt, with_particles);
        if (r) return r;
    }
    EHIPCHK(hipEventRecord(b, e->stream));
    EHIPCHK(hipEventSynchronize(b));
    float t = 0;
    EHIPCHK(hipEventElapsedTime(&t, a, b));
    *ms = t;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return BFPSH_OK;
}

int bfpsh_ensemble_time_steps_separate(bfpsh_ensemble* e, int nsteps, const double* dt, int with_particles,
                                       double* ms) {
    if (!e || !ms || nsteps < 0) return BFPSH_EINVAL;
    EHIPCHK(hipSetDevice(e->device));
    hipEvent_t a, b;
    EHIPCHK(hipEventCreate(&a));
    EHIPCHK(hipEventCreate(&b));
    EHIPCHK(hipEventRecord(a, e->stream));
    for (int i = 0; i < nsteps; i++) {
        int r = bfpsh_ensemble_step_separate(e, dt, with_particles);
        if (r) return r;
    }
    EHIPCHK(hipEventRecord(b, e->stream));
    EHIPCHK(hipEventSynchronize(b));
    float t = 0;
    EHIPCHK(hipEventElapsedTime(&t, a, b));
    *ms = t;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return BFPSH_OK;
}

int bfpsh_ensemble_synchronize(bfpsh_ensemble* e) {
    if (!e) return BFPSH_EINVAL;
    EHIPCHK(hipStreamSynchronize(e->stream));
    return BFPSH_OK;
}

int bfpsh_profile_enable(bfpsh_ctx* ctx, int enable) {
    if (!ctx) return BFPSH_EINVAL;
    int r = resolve_profile(ctx);
    if (r) return r;
    ctx->prof = enable != 0;
    if (enable) {
        ctx->prof_data.clear();
        ctx->prof_names.clear();
    }
    return BFPSH_OK;
}

int bfpsh_profile_get(bfpsh_ctx* ctx, int i, char* name, int namelen, double* total_ms, int* count) {
    if (!ctx) return BFPSH_EINVAL;
    int r = resolve_profile(ctx);
    if (r) return r;
    if (i < 0 || i >= int(ctx->prof_names.size())) return BFPSH_EINVAL;
    const std::string& s = ctx->prof_names[i];
    if (name && namelen > 0) {
        std::strncpy(name, s.c_str(), namelen - 1);
        name[namelen - 1] = 0;
    }
    const ProfEntry& pe = ctx->prof_data[s];
    if (total_ms) *total_ms = pe.ms;
    if (count) *count = pe.count;
    return BFPSH_OK;
}

int bfpsh_device_bytes(bfpsh_ctx* ctx, size_t* bytes) {
    if (!ctx || !bytes) return BFPSH_EINVAL;
    *bytes = ctx->bytes;
    return BFPSH_OK;
}

}  // extern "C"
