// Context state and launch helpers shared by the C ABI and the per-size
// kernel instantiations.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/bfpsh.h"
#include "kernels.hpp"

using namespace bfpsh;

struct ProfEntry {
    double ms = 0;
    int count = 0;
};

struct bfpsh_ctx {
    int n = 0, prec = 0, dealias = 1, device = 0;
    double dkx = 1, dky = 1, dkz = 1;
    double kM = 0, kM2 = 0, dk = 1;
    int nshells = 0;
    std::vector<double> kshell;
    std::vector<int64_t> nshell;
    hipStream_t stream = nullptr;
    bool own_stream = true;        // false: the stream of an ensemble
    void* slab = nullptr;          // ensemble member: v0|v1|v2|work(6)|ru live here
    bool force_coop = false;       // tracers: global-gather kernel, no per-step sort
    int tracer_kernel = 0;         // velocity loop and samples: 0 automatic, forced 1 bricks, 2 global gather, 3 column walk
    int tracer_form = 0;           // the form chosen for this tracer set (0: not yet)
    // device buffers (element size es = sizeof(rnumber))
    size_t es = 4;
    size_t ps = 0;                 // complex elements per plane, row pads included (plane_size)
    void* v0 = nullptr;            // cvorticity (3 planes)
    void* v1 = nullptr;            // v[1]
    void* v2 = nullptr;            // v[2]
    void* work = nullptr;          // 6 spectral planes (9 when gradients are used)
    int work_planes = 0;
    void* ru = nullptr;            // real velocity, 3 planes [y][z][x]
    void* rgrad = nullptr;         // real gradient, 9 planes
    void* tw = nullptr;            // twiddles exp(-2 pi i t / N), N complex
    double* d_famp = nullptr;      // per-stage band amplitude (fixed energy-injection forcing)
    bool ru_valid = false;
    StepParams sp{0.1, 0.0, 1.0, 2.0, 4.0, 1, 1};
    int iteration = 0;
    std::string err;
    // tracers
    int64_t np = 0;
    int pn = 0, pm = 0, nsteps = 0, step_idx = 0, head = 0, deg = 0;
    bool head_zero = false, add_head = false;
    double* d_pos = nullptr;
    double* d_rhs = nullptr;
    double* d_tab = nullptr;
    std::vector<int64_t> idx;
    std::vector<int64_t> order;    // slots in global-index order (set by particles_put)
    uint32_t* d_rank = nullptr;    // rank of each slot in global-index order
    double* d_samp = nullptr;      // sampling output, 2 x np x 9 doubles (raw, ranked)
    // standalone fields (field<rnumber, HIP, fc>): device buffer + ncomp per id
    std::vector<void*> fields;
    std::vector<int> field_nc;
    std::vector<char> field_real;  // 1 when the field holds real-space data
    // spatial ordering of the tracers (a permutation, data stays in place)
    uint32_t* d_keys = nullptr;    // 2 x np
    uint32_t* d_perm = nullptr;    // 2 x np
    void* d_sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    double* d_crec = nullptr;      // column walk: per tracer in-cell fractions (3) + packed cells (np x 4)
    uint32_t* d_bstart = nullptr;  // per-brick [start, end) (2 x nbricks)
    int64_t nbricks = 0;
    int sort_every = 1;            // re-sort period in tracer steps
    int sort_age = 1 << 30;        // steps since the last sort
    // strided-pass tile shapes: index into pass_shapes<T, N>() per pass kind
    // (BFPSH_PASS_*), chosen by the pass autotuner (bfpsh_set_plan_rigor)
    int pass_var[BFPSH_PASS_KINDS] = {0, 0, 0, 0, 0, 0};
    int plan_rigor = BFPSH_ESTIMATE;
    // profiling
    bool prof = false;
    std::vector<std::string> prof_names;
    std::map<std::string, ProfEntry> prof_data;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> ev_pool;
    size_t bytes = 0;
};

inline int fail(bfpsh_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(ctx, BFPSH_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// launch wrapper with optional per-kernel event timing: events are recorded
// around every launch on the context stream without synchronising; they are
// resolved (and recycled) by resolve_profile().
template <class F>
inline int launch(bfpsh_ctx* ctx, const char* name, F&& f) {
    hipEvent_t a = nullptr, b = nullptr;
    if (ctx->prof) {
        if (ctx->ev_pool.size() < 2) {
            for (int i = 0; i < 64; i++) {
                hipEvent_t e;
                HIPCHK(hipEventCreate(&e));
                ctx->ev_pool.push_back(e);
            }
        }
        a = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        b = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        HIPCHK(hipEventRecord(a, ctx->stream));
    }
    f();
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, BFPSH_EHIP, std::string("launch ") + name + ": " + hipGetErrorString(e));
    if (ctx->prof) {
        HIPCHK(hipEventRecord(b, ctx->stream));
        ctx->pending.push_back({name, a, b});
    }
    return BFPSH_OK;
}

inline int resolve_profile(bfpsh_ctx* ctx) {
    if (ctx->pending.empty()) return BFPSH_OK;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    for (auto& p : ctx->pending) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, p.a, p.b));
        ProfEntry& pe = ctx->prof_data[p.name];
        if (pe.count == 0) ctx->prof_names.push_back(p.name);
        pe.ms += ms;
        pe.count += 1;
        ctx->ev_pool.push_back(p.a);
        ctx->ev_pool.push_back(p.b);
    }
    ctx->pending.clear();
    return BFPSH_OK;
}

#define LAUNCH(name, ...)                                     \
    do {                                                      \
        int r_ = launch(ctx, name, [&]() { __VA_ARGS__; });   \
        if (r_) return r_;                                    \
    } while (0)

inline Geom geom(const bfpsh_ctx* ctx) {
    Geom g;
    g.n = ctx->n;
    g.dkx = ctx->dkx;
    g.dky = ctx->dky;
    g.dkz = ctx->dkz;
    g.kM2 = ctx->kM2;
    g.smooth = ctx->dealias == BFPSH_SMOOTH;
    return g;
}

// true when p lives in the context's ensemble slab (not freed by the context)
inline bool in_slab(const bfpsh_ctx* ctx, const void* p) {
    if (!ctx->slab || !p) return false;
    const char* s = (const char*)ctx->slab;
    return (const char*)p >= s && (const char*)p < s + 18 * ctx->ps * 2 * ctx->es;
}

inline int ensure_work(bfpsh_ctx* ctx, int planes) {
    if (ctx->work_planes >= planes) return BFPSH_OK;
    if (ctx->work) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (!in_slab(ctx, ctx->work)) {
            HIPCHK(hipFree(ctx->work));
            ctx->bytes -= ctx->work_planes * ctx->ps * 2 * ctx->es;
        }
    }
    HIPCHK(hipMalloc(&ctx->work, planes * ctx->ps * 2 * ctx->es));
    ctx->bytes += planes * ctx->ps * 2 * ctx->es;
    ctx->work_planes = planes;
    return BFPSH_OK;
}

inline int ensure_real(bfpsh_ctx* ctx, void** buf, int planes) {
    if (*buf) return BFPSH_OK;
    const size_t sz = size_t(planes) * ctx->ps * 2 * ctx->es;  // real planes of the padded layout
    HIPCHK(hipMalloc(buf, sz));
    HIPCHK(hipMemsetAsync(*buf, 0, sz, ctx->stream));
    ctx->bytes += sz;
    return BFPSH_OK;
}


// per-(precision, N) entry points, filled by impl_inst.hip
// A batch of equal boxes for one launch sequence: box i's v0 | v1 | v2 |
// work | ru sit i * bstride complex elements after box 0's.
struct BatchDesc {
    int nbox;
    size_t bstride;
    const StepParams* spb;  // per-box parameters (device)
    bool any_fixed;
    void *v0, *v1, *v2, *w, *ru;
};

struct OpsTable {
    int (*step)(bfpsh_ctx*, double, bool);
    int (*load_projection)(bfpsh_ctx*);
    int (*velocity_rspace)(bfpsh_ctx*);
    int (*gradient_rspace)(bfpsh_ctx*);
    int (*spectral_velocity)(bfpsh_ctx*);
    int (*cospectrum)(bfpsh_ctx*, const void*, const void*, double*);
    int (*field_ift)(bfpsh_ctx*, void*, int);
    int (*field_dft)(bfpsh_ctx*, void*, int);
    int (*batch_step)(bfpsh_ctx*, const BatchDesc&);
    int (*pressure)(bfpsh_ctx*, void*);
    int (*lagrangian_acceleration)(bfpsh_ctx*, void*);
    int (*normalize)(bfpsh_ctx*, void*, int);
    int (*rspace_stats)(bfpsh_ctx*, const void*, int, int, const double*, double*, int64_t*);
    int (*compute_vorticity)(bfpsh_ctx*);
    int (*tune_passes)(bfpsh_ctx*, int);
    PassShapes (*pass_shapes)(int kind);
};

// Process-wide pass wisdom (bfpsh.hip): the tile shape the autotuner chose
// for one (GPU architecture, precision, N, pass kind), with its time per
// plane and the rigor of the search that chose it.  Keys are
// "arch/f32|f64/N/kind".
struct PassWisdom {
    int e = 0, b = 0, pp = 0;
    double ms = 0.0;
    int rigor = BFPSH_MEASURE;  // the plan rigor that chose it
};
std::string pass_wisdom_key(const bfpsh_ctx* ctx, int kind);
bool pass_wisdom_lookup(const std::string& key, PassWisdom* w);
void pass_wisdom_store(const std::string& key, const PassWisdom& w);

int create_ctx(bfpsh_ctx** out, int nx, int ny, int nz, double dkx, double dky, double dkz, int precision,
               int dealias, int device, hipStream_t stream, void* slab);

// tracer loop and sampling (particles.hip), dispatched on precision and
// neighbour count
int bfpsh_part_loop(bfpsh_ctx* ctx, double dt);
int bfpsh_part_sample(bfpsh_ctx* ctx, const void* field, int nc, double* d_out);
int bfpsh_part_release(bfpsh_ctx* ctx);
// Tracer loops of many boxes (an ensemble, one stream) in one launch per
// neighbour count.  A PartBatch is one such group: its device table of
// boxes and their device-side Adams-Bashforth state, with host copies of
// what the device holds (so that a graph replaying the launches needs no
// host arguments, and the host rewrites them only on a change).
struct PartBatch {
    void* d = nullptr;            // device table of boxes
    void* d_dyn = nullptr;        // device AB state, 4 ints per box
    size_t cap = 0;               // table bytes allocated
    std::vector<unsigned char> tab;  // the table the device holds
    std::vector<int> dyn;            // the AB state the device holds (after the last launch)
    int nboxes = 0, nbr = 0, prec = 0, maxdeg = 0;
    uint32_t blocks = 0;          // workgroups of the loop launch
    size_t lds = 0;               // its dynamic LDS bytes
    bfpsh_ctx* ctx = nullptr;     // launches go to this box's (the ensemble's) stream
};
// Build or refresh the tables of the groups (sorting the boxes that are
// due); *ngroups receives the number of groups used.
int bfpsh_part_ens_prepare(const std::vector<bfpsh_ctx*>& boxes, const std::vector<double>& dts, PartBatch* pb,
                           int nbatch, int* ngroups);
// Launch the loops and the AB-state advance of every group (capturable).
int bfpsh_part_ens_launch(PartBatch* pb, int ngroups);
// After a launch: advance the host's view of each box and of the device state.
void bfpsh_part_ens_advanced(const std::vector<bfpsh_ctx*>& boxes, PartBatch* pb, int ngroups);
void bfpsh_part_batch_free(PartBatch& pb);

#define BFPSH_FOR_SIZES(X) X(16) X(32) X(48) X(64) X(128) X(256) X(512) X(1024)
#define BFPSH_DECL_OPS(N) const OpsTable* bfpsh_ops_f32_##N(); const OpsTable* bfpsh_ops_f64_##N();
BFPSH_FOR_SIZES(BFPSH_DECL_OPS)
#undef BFPSH_DECL_OPS
