// Launch sequences of the vorticity step, transforms and tracer loop for one
// (precision, N) pair.  Instantiated once per pair by impl_inst.hip.
#pragma once
#include <array>
#include <utility>

#include "ctx.hpp"
#include "stats.hpp"

// ------------------------------------------------------------------ per-(T, N) launchers
template <typename T, int N>
struct Impl {
    using CF = Cfg<T, N>;
    static constexpr int B = CF::B, NT = CF::NT_PASS, NXH = CF::NXH;
    typedef C2<T> CT;

    // A batch of equal boxes (same N, precision, dk, dealiasing) whose
    // v0 | v1 | v2 | work | ru buffers sit bstride complex elements apart;
    // one context is a batch of one with bstride = 0.
    struct Batch {
        int nbox = 1;
        size_t bstride = 0;
        const StepParams* spb = nullptr;  // per-box parameters (device), or null
        bool any_fixed = false;           // some box uses the fixed-injection forcing
        CT *v0, *v1, *v2, *w;
        T* ru;
        double* famp = nullptr;           // single box: its 3 band amplitudes
        StepParams sp;
        Geom g;
    };
    // ---------------------------------------------------------- strided passes
    // compiled tile shapes of k_pass for this (T, N); ctx->pass_var[kind]
    // selects one per pass kind
    static constexpr PassShapes SH = pass_shapes<T, N>();
    typedef void (*PassFn)(hipStream_t, CT*, size_t, size_t, const CT*, int, size_t, int);
    template <int SIGN, int I>
    static void pass_launch(hipStream_t st, CT* d, size_t es, size_t os, const CT* tw, int planes, size_t bs,
                            int nbox) {
        constexpr PassShape p = SH.v[I < SH.n ? I : 0];
        constexpr int nt = (N / p.e) * p.b;
        if constexpr (p.pp) {
            // as many workgroups as are resident at once (the occupancy of
            // this kernel on this device, looked up once)
            static int grid = 0;
            if (grid == 0) {
                int dev = 0, ncu = 0, occ = 0;
                (void)hipGetDevice(&dev);
                (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
                (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_pass_pp<T, N, SIGN, p.e, p.b>, nt, 0);
                grid = std::max(1, ncu) * std::max(1, occ);
            }
            const int ntx = NXH / p.b, nty = N, ntiles = ntx * nty * planes * nbox;
            k_pass_pp<T, N, SIGN, p.e, p.b><<<std::min(grid, ntiles), nt, 0, st>>>(d, es, os, tw, planes, bs, ntx,
                                                                                   nty, ntiles);
        } else {
            k_pass<T, N, SIGN, p.e, p.b><<<dim3(NXH / p.b, N, planes * nbox), nt, 0, st>>>(d, es, os, tw, planes, bs);
        }
    }
    template <int SIGN, size_t... I>
    static constexpr std::array<PassFn, kMaxPassShapes> pass_table(std::index_sequence<I...>) {
        return {{&pass_launch<SIGN, int(I)>...}};
    }
    static PassFn pass_fn(int kind, int v) {
        static constexpr std::array<PassFn, kMaxPassShapes> inv = pass_table<+1>(std::make_index_sequence<kMaxPassShapes>{});
        static constexpr std::array<PassFn, kMaxPassShapes> fwd = pass_table<-1>(std::make_index_sequence<kMaxPassShapes>{});
        if (v < 0 || v >= SH.n) v = 0;
        return (kind == BFPSH_PASS_Z_INV || kind == BFPSH_PASS_Y_INV) ? inv[v] : fwd[v];
    }
    // element and tile strides of a pass kind: y passes walk ky rows, z
    // passes the kz rows of one ky slab
    static size_t pass_estride(int kind) {
        return (kind == BFPSH_PASS_Y_INV || kind == BFPSH_PASS_Y_FWD) ? row_stride(N) : size_t(NXH);
    }
    static size_t pass_ostride(int kind) {
        return (kind == BFPSH_PASS_Y_INV || kind == BFPSH_PASS_Y_FWD) ? size_t(NXH) : row_stride(N);
    }
    // one in-place strided pass of `planes` planes of each of nbox boxes
    // (bstride complex elements apart)
    static int pass(bfpsh_ctx* ctx, int kind, CT* d, int planes, const char* name, int nbox = 1,
                    size_t bstride = 0) {
        const PassFn f = pass_fn(kind, ctx->pass_var[kind]);
        LAUNCH(name, f(ctx->stream, d, pass_estride(kind), pass_ostride(kind), (const CT*)ctx->tw, planes, bstride,
                       nbox));
        return BFPSH_OK;
    }
    static PassShapes shapes() { return SH; }

    // ---------------------------------------------------------- fused passes
    // tile shapes of the Biot-Savart first pass (MODE 0, the step's) and of
    // the last forward pass with the update; ctx->pass_var[BFPSH_PASS_BS]
    // and [BFPSH_PASS_Y3] select one (the velocity and gradient forms of the
    // first pass and the batched forms keep the built-in shape)
    static constexpr PassShapes SH_BS = fused_shapes<T, N>(CF::E_BS, CF::B_BS);
    static constexpr PassShapes SH_Y3 = fused_shapes<T, N>(CF::E_Y3, CF::B_Y3);
    typedef void (*BsFn)(hipStream_t, const CT*, CT*, const CT*, const Geom&, int, size_t);
    template <int I>
    static void bs_launch(hipStream_t st, const CT* src, CT* w, const CT* tw, const Geom& g, int nbox, size_t bstride) {
        constexpr PassShape p = SH_BS.v[I < SH_BS.n ? I : 0];
        constexpr int nt = (N / p.e) * p.b;
        k_inv_z_bs<T, N, 0, true, p.e, p.b><<<dim3(1, N, nbox), nt, 0, st>>>(src, w, tw, g, 0, bstride);
        if (NXH / p.b > 1)
            k_inv_z_bs<T, N, 0, false, p.e, p.b><<<dim3(NXH / p.b - 1, N, nbox), nt, 0, st>>>(src, w, tw, g, 1,
                                                                                              bstride);
    }
    template <size_t... I>
    static constexpr std::array<BsFn, kMaxPassShapes> bs_table(std::index_sequence<I...>) {
        return {{&bs_launch<int(I)>...}};
    }
    static BsFn bs_fn(int v) {
        static constexpr std::array<BsFn, kMaxPassShapes> t = bs_table(std::make_index_sequence<kMaxPassShapes>{});
        return t[(v >= 0 && v < SH_BS.n) ? v : 0];
    }
    typedef void (*Y3Fn)(hipStream_t, const CT*, const CT*, const CT*, CT*, const CT*, const Geom&, const StepParams&);
    template <int S, int I>
    static void y3_launch(hipStream_t st, const CT* w, const CT* v0, const CT* vp, CT* dst, const CT* tw,
                          const Geom& g, const StepParams& sp) {
        constexpr PassShape p = SH_Y3.v[I < SH_Y3.n ? I : 0];
        k_fwd_y_update<T, N, S, false, p.e, p.b><<<dim3(NXH / p.b, N, 1), (N / p.e) * p.b, 0, st>>>(
            w, v0, vp, dst, tw, g, sp, 0, nullptr);
    }
    template <int S, size_t... I>
    static constexpr std::array<Y3Fn, kMaxPassShapes> y3_table(std::index_sequence<I...>) {
        return {{&y3_launch<S, int(I)>...}};
    }
    static Y3Fn y3_fn(int s, int v) {
        static constexpr std::array<Y3Fn, kMaxPassShapes> t1 = y3_table<1>(std::make_index_sequence<kMaxPassShapes>{});
        static constexpr std::array<Y3Fn, kMaxPassShapes> t2 = y3_table<2>(std::make_index_sequence<kMaxPassShapes>{});
        static constexpr std::array<Y3Fn, kMaxPassShapes> t3 = y3_table<3>(std::make_index_sequence<kMaxPassShapes>{});
        v = (v >= 0 && v < SH_Y3.n) ? v : 0;
        return s == 1 ? t1[v] : (s == 2 ? t2[v] : t3[v]);
    }
    static PassShapes shapes_of(int kind) {
        return kind == BFPSH_PASS_BS ? SH_BS : (kind == BFPSH_PASS_Y3 ? SH_Y3 : SH);
    }

    // The pass autotuner: for each pass kind, the shape held by the wisdom
    // when a search of at least this rigor chose it (any wisdom for
    // ESTIMATE), else the fastest of all compiled shapes timed on one
    // scratch plane (ESTIMATE without wisdom: the built-in shape), as FFTW's
    // planner searches its algorithms under the reference's plan rigor
    // (bfps/cpp/fftw_interface.hpp:30-35) and reuses its wisdom.
    static int tune_passes(bfpsh_ctx* ctx, int rigor) {
        ctx->plan_rigor = rigor;
        {
            int r = ensure_work(ctx, 6);
            if (r) return r;
        }
        void* scratch = nullptr;
        hipEvent_t a = nullptr, b = nullptr;
        int rc = BFPSH_OK;
        for (int kind = 0; kind < BFPSH_PASS_KINDS && rc == BFPSH_OK; kind++) {
            ctx->pass_var[kind] = 0;
            const PassShapes sh = shapes_of(kind);
            if (sh.n == 1) continue;
            const std::string key = pass_wisdom_key(ctx, kind);
            PassWisdom w;
            if (pass_wisdom_lookup(key, &w) && w.rigor >= rigor) {
                for (int v = 0; v < sh.n; v++)
                    if (sh.v[v].e == w.e && sh.v[v].b == w.b && sh.v[v].pp == w.pp) ctx->pass_var[kind] = v;
                continue;
            }
            if (rigor == BFPSH_ESTIMATE) continue;
            if (!scratch) {
                const size_t bytes = ctx->ps * sizeof(CT);
                if (hipMalloc(&scratch, bytes) != hipSuccess || hipMemsetAsync(scratch, 0, bytes, ctx->stream) != hipSuccess ||
                    hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
                    rc = fail(ctx, BFPSH_EHIP, "pass autotuner: scratch plane or events");
                    break;
                }
            }
            const int reps = rigor >= BFPSH_PATIENT ? 10 : 3;
            double best = 1e300;
            int bestv = 0;
            for (int v = 0; v < sh.n && rc == BFPSH_OK; v++) {
                const CT* tw = (const CT*)ctx->tw;
                // one launch of shape v: a plain pass over the scratch plane;
                // the first pass from cvorticity into the work planes; the
                // last pass (stage 2) from work planes 0..2 into 3..5
                auto run = [&]() {
                    if (kind == BFPSH_PASS_BS) {
                        bs_fn(v)(ctx->stream, (const CT*)ctx->v0, (CT*)ctx->work, tw, geom(ctx), 1, 0);
                    } else if (kind == BFPSH_PASS_Y3) {
                        StepParams sp = ctx->sp;
                        sp.forcing = BFPSH_FORCING_NONE;
                        sp.famp_dev = nullptr;
                        const CT* w = (const CT*)ctx->work;
                        y3_fn(2, v)(ctx->stream, w, (const CT*)ctx->v0, (const CT*)ctx->v0,
                                    (CT*)ctx->work + 3 * ctx->ps, tw, geom(ctx), sp);
                    } else {
                        pass_fn(kind, v)(ctx->stream, static_cast<CT*>(scratch), pass_estride(kind),
                                         pass_ostride(kind), tw, 1, 0, 1);
                    }
                };
                run();  // warm-up
                (void)hipEventRecord(a, ctx->stream);
                for (int r = 0; r < reps; r++) run();
                (void)hipEventRecord(b, ctx->stream);
                hipError_t e = hipGetLastError();
                if (e == hipSuccess) e = hipEventSynchronize(b);
                float ms = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
                if (e != hipSuccess) {
                    rc = fail(ctx, BFPSH_EHIP, std::string("pass autotuner: ") + hipGetErrorString(e));
                    break;
                }
                if (ms / reps < best) {
                    best = ms / reps;
                    bestv = v;
                }
            }
            if (rc == BFPSH_OK) {
                ctx->pass_var[kind] = bestv;
                w.e = sh.v[bestv].e;
                w.b = sh.v[bestv].b;
                w.pp = sh.v[bestv].pp;
                w.ms = best;
                w.rigor = rigor;
                pass_wisdom_store(key, w);
            }
        }
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        if (scratch) {
            (void)hipStreamSynchronize(ctx->stream);
            (void)hipFree(scratch);
        }
        return rc;
    }

    static Batch single(bfpsh_ctx* ctx, double dt) {
        Batch b;
        b.v0 = (CT*)ctx->v0;
        b.v1 = (CT*)ctx->v1;
        b.v2 = (CT*)ctx->v2;
        b.w = (CT*)ctx->work;
        b.ru = nullptr;
        b.sp = ctx->sp;
        b.sp.dt = dt;
        b.famp = ctx->d_famp;
        b.sp.famp_dev = ctx->d_famp;
        b.any_fixed = b.sp.forcing == BFPSH_FORCING_FIXED_ENERGY;
        b.g = geom(ctx);
        return b;
    }

    // Biot-Savart + first inverse pass: kx tile 0 with the symmetrize
    // substitution, the remaining tiles without
    template <int MODE>
    static int bs_pass(bfpsh_ctx* ctx, const CT* src, CT* w, const CT* tw, const Geom& g, const char* name,
                       int nbox = 1, size_t bstride = 0) {
        LAUNCH("inv_z_bs_t0", (k_inv_z_bs<T, N, MODE, true><<<dim3(1, N, nbox), CF::NT_BS, 0, ctx->stream>>>(
                                  src, w, tw, g, 0, bstride)));
        if (NXH / CF::B_BS > 1)
            LAUNCH(name, (k_inv_z_bs<T, N, MODE, false><<<dim3(NXH / CF::B_BS - 1, N, nbox), CF::NT_BS, 0, ctx->stream>>>(
                              src, w, tw, g, 1, bstride)));
        return BFPSH_OK;
    }

    // RK stage s of every box of the batch; ru (box 0 address) receives the
    // real velocity of the stage-1 source when non-null
    static int stage(bfpsh_ctx* ctx, int s, const Batch& bt, T* ru) {
        const Geom& g = bt.g;
        const int nb = bt.nbox;
        const size_t bs = bt.bstride;
        CT* v0 = bt.v0;
        CT* v1 = bt.v1;
        CT* v2 = bt.v2;
        CT* w = bt.w;
        const CT* tw = (const CT*)ctx->tw;
        const CT* src = s == 1 ? v0 : (s == 2 ? v1 : v2);
        if (bt.any_fixed) {
            if (bt.spb)
                LAUNCH("band_amplitude", (k_band_amplitude<T, N><<<nb, 256, 0, ctx->stream>>>(
                                             src, nullptr, g, 0.0, 0.0, 0.0, bs, bt.spb, s)));
            else
                LAUNCH("band_amplitude", (k_band_amplitude<T, N><<<1, 256, 0, ctx->stream>>>(
                                             src, bt.famp + (s - 1), g, bt.sp.famp, bt.sp.fk0, bt.sp.fk1, 0, nullptr, s)));
        }
        if (bt.spb) {
            int r = bs_pass<0>(ctx, src, w, tw, g, "inv_z_bs", nb, bs);
            if (r) return r;
        } else {
            LAUNCH("inv_z_bs", bs_fn(ctx->pass_var[BFPSH_PASS_BS])(ctx->stream, src, w, tw, g, 1, 0));
        }
        {
            int r = pass(ctx, BFPSH_PASS_Y_INV, w, 6, "inv_y", nb, bs);
            if (r) return r;
        }
        if constexpr (RowsW<T, N>::ok && !BFPSH_ROWS_TILE)
            LAUNCH("rows_cross", (k_rows_cross_w<T, N><<<dim3(N * N / 2 / RowsW<T, N>::RPB, nb), RowsW<T, N>::NT, 0,
                                                         ctx->stream>>>(w, ru, tw, T(double(N) * N * N), bs)));
        else
            LAUNCH("rows_cross", (k_rows_cross<T, N><<<dim3(N * N / 2 / CF::RP, nb), CF::NT_ROW, 0, ctx->stream>>>(
                                      w, ru, tw, T(double(N) * N * N), bs)));
        {
            int r = pass(ctx, BFPSH_PASS_Z_FWD, w, 3, "fwd_z", nb, bs);
            if (r) return r;
        }
        const dim3 tiles3(NXH / CF::B_Y3, N, nb);
        const char* nm = s == 1 ? "fwd_y_update1" : (s == 2 ? "fwd_y_update2" : "fwd_y_update3");
        const CT* vp = s == 1 ? v1 : (s == 2 ? v1 : v2);
        CT* dst = s == 1 ? v1 : (s == 2 ? v2 : v0);
        if (bt.spb) {
            if (s == 1) LAUNCH(nm, (k_fwd_y_update<T, N, 1, true><<<tiles3, CF::NT_Y3, 0, ctx->stream>>>(w, v0, vp, dst, tw, g, bt.sp, bs, bt.spb)));
            else if (s == 2) LAUNCH(nm, (k_fwd_y_update<T, N, 2, true><<<tiles3, CF::NT_Y3, 0, ctx->stream>>>(w, v0, vp, dst, tw, g, bt.sp, bs, bt.spb)));
            else LAUNCH(nm, (k_fwd_y_update<T, N, 3, true><<<tiles3, CF::NT_Y3, 0, ctx->stream>>>(w, v0, vp, dst, tw, g, bt.sp, bs, bt.spb)));
        } else {
            LAUNCH(nm, y3_fn(s, ctx->pass_var[BFPSH_PASS_Y3])(ctx->stream, w, v0, vp, dst, tw, g, bt.sp));
        }
        return BFPSH_OK;
    }

    static int step(bfpsh_ctx* ctx, double dt, bool particles) {
        int r = ensure_work(ctx, 6);
        if (r) return r;
        T* ru = nullptr;
        if (particles) {
            r = ensure_real(ctx, &ctx->ru, 3);
            if (r) return r;
            ru = (T*)ctx->ru;
        }
        const Batch bt = single(ctx, dt);
        r = stage(ctx, 1, bt, ru);
        if (r) return r;
        if (particles) {
            r = bfpsh_part_loop(ctx, dt);
            if (r) return r;
        }
        r = stage(ctx, 2, bt, nullptr);
        if (r) return r;
        r = stage(ctx, 3, bt, nullptr);
        if (r) return r;
        r = symmetrize(ctx, ctx->v0, 3);
        if (r) return r;
        ctx->ru_valid = false;
        ctx->iteration++;
        return BFPSH_OK;
    }

    static int batch_step_desc(bfpsh_ctx* ctx, const BatchDesc& d) {
        Batch bt;
        bt.nbox = d.nbox;
        bt.bstride = d.bstride;
        bt.spb = d.spb;
        bt.any_fixed = d.any_fixed;
        bt.v0 = (CT*)d.v0;
        bt.v1 = (CT*)d.v1;
        bt.v2 = (CT*)d.v2;
        bt.w = (CT*)d.w;
        bt.ru = (T*)d.ru;
        bt.sp = ctx->sp;
        bt.g = geom(ctx);
        return batch_step(ctx, bt);
    }

    // The field part of one step of a whole batch (the tracers of each box
    // are advanced separately with the stage-1 real velocity left in ru)
    static int batch_step(bfpsh_ctx* ctx, const Batch& bt) {
        for (int s = 1; s <= 3; s++) {
            int r = stage(ctx, s, bt, s == 1 ? bt.ru : nullptr);
            if (r) return r;
        }
        const int nthreads = 256;
        LAUNCH("symmetrize", (k_symmetrize<T><<<dim3((N * N + nthreads - 1) / nthreads, bt.nbox), nthreads, 0,
                                                 ctx->stream>>>(bt.v0, N, 3, bt.bstride)));
        return BFPSH_OK;
    }

    static int symmetrize(bfpsh_ctx* ctx, void* v, int nc) {
        const int nthreads = 256;
        const int nblocks = (N * N + nthreads - 1) / nthreads;
        LAUNCH("symmetrize", (k_symmetrize<T><<<nblocks, nthreads, 0, ctx->stream>>>((CT*)v, N, nc, 0)));
        return BFPSH_OK;
    }

    static int load_projection(bfpsh_ctx* ctx) {
        int r = symmetrize(ctx, ctx->v0, 3);
        if (r) return r;
        const size_t ps = size_t(N) * N * NXH;
        LAUNCH("lowpass_divfree", (k_lowpass_divfree<T><<<(ps + 255) / 256, 256, 0, ctx->stream>>>((CT*)ctx->v0, geom(ctx))));
        ctx->ru_valid = false;
        return BFPSH_OK;
    }

    static int velocity_rspace(bfpsh_ctx* ctx) {
        int r = ensure_work(ctx, 6);
        if (r) return r;
        r = ensure_real(ctx, &ctx->ru, 3);
        if (r) return r;
        CT* w = (CT*)ctx->work;
        const CT* tw = (const CT*)ctx->tw;
        r = bs_pass<1>(ctx, (const CT*)ctx->v0, w, tw, geom(ctx), "inv_z_bs_u");
        if (r) return r;
        r = pass(ctx, BFPSH_PASS_Y_INV, w, 3, "inv_y");
        if (r) return r;
        LAUNCH("rows_c2r3", (k_rows_c2r<T, N, 3><<<N * N / 2 / CF::RP, CF::NT_ROW, 0, ctx->stream>>>(w, (T*)ctx->ru, tw)));
        ctx->ru_valid = true;
        return BFPSH_OK;
    }

    static int gradient_rspace(bfpsh_ctx* ctx) {
        int r = ensure_work(ctx, 9);
        if (r) return r;
        r = ensure_real(ctx, &ctx->rgrad, 9);
        if (r) return r;
        CT* w = (CT*)ctx->work;
        const CT* tw = (const CT*)ctx->tw;
        r = bs_pass<2>(ctx, (const CT*)ctx->v0, w, tw, geom(ctx), "inv_z_bs_grad");
        if (r) return r;
        r = pass(ctx, BFPSH_PASS_Y_INV, w, 9, "inv_y");
        if (r) return r;
        LAUNCH("rows_c2r9", (k_rows_c2r<T, N, 9><<<N * N / 2 / CF::RP, CF::NT_ROW, 0, ctx->stream>>>(w, (T*)ctx->rgrad, tw)));
        return BFPSH_OK;
    }

    // field::ift / field::dft on a standalone NC-component field (in place):
    // c2r = z pass, y pass (inverse), x rows; r2c = x rows, z pass, y pass
    template <int NC>
    static int field_ift_nc(bfpsh_ctx* ctx, void* data) {
        CT* d = (CT*)data;
        const CT* tw = (const CT*)ctx->tw;
        int r = pass(ctx, BFPSH_PASS_Z_INV, d, NC, "field_inv_z");
        if (r) return r;
        r = pass(ctx, BFPSH_PASS_Y_INV, d, NC, "field_inv_y");
        if (r) return r;
        LAUNCH("field_rows_c2r", (k_rows_c2r<T, N, NC><<<N * N / 2 / CF::RP, CF::NT_ROW, 0, ctx->stream>>>(d, (T*)d, tw)));
        return BFPSH_OK;
    }
    template <int NC>
    static int field_dft_nc(bfpsh_ctx* ctx, void* data) {
        CT* d = (CT*)data;
        const CT* tw = (const CT*)ctx->tw;
        LAUNCH("field_rows_r2c", (k_rows_r2c<T, N, NC><<<N * N / 2 / CF::RP, CF::NT_ROW, 0, ctx->stream>>>((const T*)d, d, tw)));
        int r = pass(ctx, BFPSH_PASS_Z_FWD, d, NC, "field_fwd_z");
        if (r) return r;
        r = pass(ctx, BFPSH_PASS_Y_FWD, d, NC, "field_fwd_y");
        if (r) return r;
        return BFPSH_OK;
    }
    static int field_ift(bfpsh_ctx* ctx, void* data, int nc) {
        return nc == 1 ? field_ift_nc<1>(ctx, data) : (nc == 3 ? field_ift_nc<3>(ctx, data) : field_ift_nc<9>(ctx, data));
    }
    static int field_dft(bfpsh_ctx* ctx, void* data, int nc) {
        return nc == 1 ? field_dft_nc<1>(ctx, data) : (nc == 3 ? field_dft_nc<3>(ctx, data) : field_dft_nc<9>(ctx, data));
    }

    // cvelocity = symmetrize(BS(cvorticity)) into work planes 0..2
    static int spectral_velocity(bfpsh_ctx* ctx) {
        int r = ensure_work(ctx, 6);
        if (r) return r;
        const size_t ps = size_t(N) * N * NXH;
        LAUNCH("biot_savart", (k_biot_savart<T><<<(ps + 255) / 256, 256, 0, ctx->stream>>>((const CT*)ctx->v0, (CT*)ctx->work, geom(ctx))));
        return symmetrize(ctx, ctx->work, 3);
    }

    // compute_vorticity (vorticity_equation.cpp:171-204): cvorticity =
    // symmetrize(i k x cvelocity), cvelocity in work planes 0..2
    static int compute_vorticity(bfpsh_ctx* ctx) {
        int r = ensure_work(ctx, 6);
        if (r) return r;
        const size_t ps = size_t(N) * N * NXH;
        LAUNCH("curl", (k_curl<T><<<(ps + 255) / 256, 256, 0, ctx->stream>>>((const CT*)ctx->work, (CT*)ctx->v0, geom(ctx))));
        ctx->ru_valid = false;
        return symmetrize(ctx, ctx->v0, 3);
    }

    // compute_pressure (vorticity_equation.cpp:441-520) of the velocity of
    // cvorticity into the spectral plane p: the real velocity is made here
    // (the reference takes `u` in real space), its products go through work
    // planes 0..2
    static int pressure(bfpsh_ctx* ctx, CT* p) {
        int r = velocity_rspace(ctx);
        if (r) return r;
        // products over whole real planes (pads included), k over the logical half spectrum
        const size_t rpn = 2 * plane_size(N), ps = size_t(N) * N * NXH;
        const unsigned gr = unsigned(std::min<size_t>((rpn + 255) / 256, 65536));
        for (int pass = 0; pass < 2; pass++) {
            LAUNCH("pressure_uu", (k_uu<T><<<gr, 256, 0, ctx->stream>>>((const T*)ctx->ru, (T*)ctx->work, rpn, pass)));
            r = field_dft_nc<3>(ctx, ctx->work);
            if (r) return r;
            LAUNCH("pressure", (k_pressure<T><<<unsigned((ps + 255) / 256), 256, 0, ctx->stream>>>(
                                   (const CT*)ctx->work, p, geom(ctx), pass)));
        }
        return BFPSH_OK;
    }

    // compute_Lagrangian_acceleration (vorticity_equation.cpp:522-578) into
    // the 3 spectral planes acc; the pressure lives in work plane 3
    static int lagrangian_acceleration(bfpsh_ctx* ctx, CT* acc) {
        int r = ensure_work(ctx, 6);
        if (r) return r;
        const size_t ps = size_t(N) * N * NXH;  // logical modes per plane
        CT* p = (CT*)ctx->work + 3 * plane_size(N);
        r = pressure(ctx, p);
        if (r) return r;
        r = spectral_velocity(ctx);
        if (r) return r;
        const StepParams& sp = ctx->sp;
        LAUNCH("lagrangian_acceleration", (k_lagr_acc<T><<<unsigned((ps + 255) / 256), 256, 0, ctx->stream>>>(
                                              (const CT*)ctx->work, p, acc, geom(ctx), sp.nu, sp.forcing == 1 ? 1 : 0,
                                              sp.famp, sp.fk0, sp.fk1)));
        return BFPSH_OK;
    }

    static int pressure_v(bfpsh_ctx* ctx, void* p) { return pressure(ctx, (CT*)p); }
    static int lagrangian_acceleration_v(bfpsh_ctx* ctx, void* a) { return lagrangian_acceleration(ctx, (CT*)a); }

    // field::normalize of an NC-component field (all its values; the row
    // pads are divided too and never read)
    static int normalize(bfpsh_ctx* ctx, void* data, int nc) {
        const size_t cnt = size_t(nc) * 2 * plane_size(N);
        const unsigned gr = unsigned(std::min<size_t>((cnt + 255) / 256, 65536));
        LAUNCH("normalize", (k_div<T><<<gr, 256, 0, ctx->stream>>>((T*)data, cnt, T(double(N) * N * N))));
        return BFPSH_OK;
    }

    // field::compute_rspace_stats (field.cpp:661-850) of a real NC-component
    // field: moments [10][nvals], hist [nbins][nvals]
    static int rspace_stats(bfpsh_ctx* ctx, const void* data, int nc, int nbins, const double* maxest,
                            double* moments, int64_t* hist) {
        const size_t n3 = size_t(N) * N * N;
        const int nvals = nc == 3 ? 4 : nc, nv = nc == 3 ? 4 : 1, groups = nc == 3 ? 1 : nc;
        const int nblk = int(std::min<size_t>(1024, (n3 + 255) / 256));
        const size_t part_n = size_t(groups) * nblk * 10 * nv;
        double *d_me = nullptr, *d_part = nullptr, *d_mom = nullptr;
        unsigned long long* d_hist = nullptr;
        HIPCHK(hipMallocAsync((void**)&d_me, nvals * sizeof(double), ctx->stream));
        HIPCHK(hipMallocAsync((void**)&d_part, part_n * sizeof(double), ctx->stream));
        HIPCHK(hipMallocAsync((void**)&d_mom, 10 * nvals * sizeof(double), ctx->stream));
        HIPCHK(hipMallocAsync((void**)&d_hist, size_t(nbins) * nvals * sizeof(unsigned long long), ctx->stream));
        HIPCHK(hipMemcpyAsync(d_me, maxest, nvals * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemsetAsync(d_hist, 0, size_t(nbins) * nvals * sizeof(unsigned long long), ctx->stream));
        const size_t lds = size_t(nbins) * nv * sizeof(unsigned int);
        if (nc == 3) {
            LAUNCH("rspace_stats", (k_rstats<T, 4><<<dim3(nblk, 1), 256, lds, ctx->stream>>>(
                                       (const T*)data, N, nvals, nbins, d_me, d_part, d_hist)));
            LAUNCH("rspace_stats_final", (k_rstats_final<4><<<1, 64, 0, ctx->stream>>>(d_part, nblk, nvals,
                                                                                       double(n3), d_mom)));
        } else {
            LAUNCH("rspace_stats", (k_rstats<T, 1><<<dim3(nblk, nc), 256, lds, ctx->stream>>>(
                                       (const T*)data, N, nvals, nbins, d_me, d_part, d_hist)));
            LAUNCH("rspace_stats_final", (k_rstats_final<1><<<nc, 64, 0, ctx->stream>>>(d_part, nblk, nvals,
                                                                                        double(n3), d_mom)));
        }
        HIPCHK(hipMemcpyAsync(moments, d_mom, 10 * nvals * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(hist, d_hist, size_t(nbins) * nvals * sizeof(int64_t), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipFreeAsync(d_me, ctx->stream));
        HIPCHK(hipFreeAsync(d_part, ctx->stream));
        HIPCHK(hipFreeAsync(d_mom, ctx->stream));
        HIPCHK(hipFreeAsync(d_hist, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return BFPSH_OK;
    }

    static int cospectrum(bfpsh_ctx* ctx, const void* a, const void* b, double* out) {
        double* d_out = nullptr;
        HIPCHK(hipMallocAsync((void**)&d_out, ctx->nshells * 9 * sizeof(double), ctx->stream));
        HIPCHK(hipMemsetAsync(d_out, 0, ctx->nshells * 9 * sizeof(double), ctx->stream));
        const size_t lds = ctx->nshells * 9 * sizeof(double);
        if (lds > 64 * 1024) return fail(ctx, BFPSH_EINVAL, "too many shells for the cospectrum kernel");
        LAUNCH("cospectrum", (k_cospectrum<T><<<1024, 256, lds, ctx->stream>>>((const CT*)a, (const CT*)b, d_out, geom(ctx), ctx->dk, ctx->nshells)));
        HIPCHK(hipMemcpyAsync(out, d_out, ctx->nshells * 9 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipFreeAsync(d_out, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return BFPSH_OK;
    }

};

