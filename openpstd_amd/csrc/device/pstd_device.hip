// MI355X (gfx950) device solver for openPSTD's PSTD update loop.
//
// Implements the pstd_ctx_* C ABI of include/openpstd_amd.h: the GPU solver
// slot of openPSTD (GPUSingleThreadSolver / GPUMultiThreadSolver,
// kernel/Solver.cpp:153-160).
//
// State per domain lives in four arenas with identical layout (fields vx, vy,
// px, py row-major at per-domain offsets):
//   TN  - values at t_n after PML attenuation (openPSTD's previous_values)
//   S0, S1 - raw outputs of the RK stages (ping-pong)
//   P0I - an explicit p0 (initial state / host-written p0)
// p0 is never stored between stages: openPSTD recomputes p0 = px0 + py0 after
// every stage (kernel/Solver.cpp:114-116), so the derivative kernels form it on
// the fly from the stage outputs.  The first stage of a step reads p0 from the
// raw stage-5 outputs of the previous step, which is openPSTD's p0 left stale
// by the PML attenuation (kernel/Solver.cpp:85).
//
// One RK stage = one launch of the fused line kernel: for each line of each
// segment it gathers the windowed line (spatderp3, kernel/core/
// kernel_functions.cpp:191-307), applies the spectral derivative with two
// complex FFTs in LDS (two real lines per complex transform), and writes the
// stage update of Solver::update_field_values (kernel/Solver.cpp:174-181).  The
// last stage also writes TN = output * PML factor (Domain::apply_pml_matrices,
// kernel/core/Domain.cpp:716-725).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../../include/openpstd_amd.h"
#include "fft_line.h"

using namespace pstdamd::dev;

namespace {

constexpr int kSmallThreads = 256;  // block size of the F <= 2048 class
constexpr int kPassChunk = 2048;    // elements per passthrough tile

struct DomainDev {
    int nx, ny;
    int off_vx, off_vy, off_px, off_py;
    float rho;
    int is_pml;
};

struct SegDev {
    int domain, dir, field;
    int side1, side2;
    int off1, off2;
    int wlen;
    int n_main;    // main line length (n for pressure, n+1 for velocity)
    int n_out;     // derivative values per line (n+1 pressure, n velocity)
    int F;
    int h_off;     // float2 offset of the derivative filter in the table buffer
    int tw_off;    // float2 offset of the twiddle table
    int win_off;   // float offset of the 2w+1 window
    float c_s1, c_m1, c_s2, c_m2;
};

struct TileDev {
    int seg;    // derivative segment, or -1 for an elementwise passthrough tile
    int line0;  // first local line (derivative) / first element (passthrough)
    int count;  // lines (derivative) / elements (passthrough)
    int domain; // passthrough: domain
    int field;  // passthrough: PSTD_FIELD_*
};

struct StageArgs {
    const DomainDev *doms;
    const SegDev *segs;
    const TileDev *tiles;
    int ntiles;
    const float2 *ctab;
    const float *ftab;
    const float *p0a;  // p0 = p0a[px slot] + (p0b ? p0b[py slot] : 0)
    const float *p0b;
    const float *vsrc; // current vx / vy
    float *tn;
    float *out;
    const float *pml;  // PML factor arena (same layout); 1 for non-PML domains
    float s;           // dt * alpha_k
    float c2;          // c1 * c1
    int last;          // last RK stage: also write tn = out * pml
};

// ------------------------------------------------------------- device code

__device__ __forceinline__ int field_offset(const DomainDev &d, int f) {
    return f == PSTD_FIELD_VX ? d.off_vx : f == PSTD_FIELD_VY ? d.off_vy : f == PSTD_FIELD_PX ? d.off_px : d.off_py;
}

// Value of a domain's derivative source at (line, pos): p0 for pressure
// segments, vx (dir X) / vy (dir Y) for velocity segments.
__device__ __forceinline__ float src_value(const StageArgs &a, const DomainDev &d, int dir, int field, int line,
                                           int pos) {
    if (field == 0) {
        int idx = dir == 0 ? line * d.nx + pos : pos * d.nx + line;
        float v = a.p0a[d.off_px + idx];
        if (a.p0b) v += a.p0b[d.off_py + idx];
        return v;
    }
    if (dir == 0) return a.vsrc[d.off_vx + line * (d.nx + 1) + pos];
    return a.vsrc[d.off_vy + pos * d.nx + line];
}

// Element i of the windowed line [dom1 | main | dom3 | 0] (spatderp3,
// kernel/core/kernel_functions.cpp:210-216 pressure, :263-269 velocity).
__device__ __forceinline__ float line_value(const StageArgs &a, const SegDev &sg, const DomainDev &dm, int line,
                                            int i) {
    const int w = sg.wlen;
    const int nm = sg.n_main;
    if (i >= w && i < w + nm) return src_value(a, dm, sg.dir, sg.field, line, i - w);
    const bool vel = sg.field == 1;
    if (i < w) {
        // a missing neighbour is a zero block with transparent coefficients: 0
        if (sg.side1 < 0) return 0.f;
        const int j = i;
        const DomainDev &d1 = a.doms[sg.side1];
        const int s1len = (sg.dir == 0 ? d1.nx : d1.ny) + (vel ? 1 : 0);
        float side = src_value(a, d1, sg.dir, sg.field, line + sg.off1, vel ? s1len - w - 1 + j : s1len - w + j);
        float main = src_value(a, dm, sg.dir, sg.field, line, vel ? w - j : w - 1 - j);
        return (side * sg.c_s1 + main * sg.c_m1) * a.ftab[sg.win_off + j];
    }
    if (i < 2 * w + nm) {
        const int j = i - w - nm;
        if (sg.side2 < 0) return 0.f;
        const DomainDev &d2 = a.doms[sg.side2];
        float main = src_value(a, dm, sg.dir, sg.field, line, vel ? nm - 2 - j : nm - 1 - j);
        float side = src_value(a, d2, sg.dir, sg.field, line + sg.off2, vel ? 1 + j : j);
        return (side * sg.c_s2 + main * sg.c_m2) * a.ftab[sg.win_off + w + 1 + j];
    }
    return 0.f;
}

// Solver::update_field_values (kernel/Solver.cpp:174-181) for one value.
__device__ __forceinline__ void write_update(const StageArgs &a, const SegDev &sg, const DomainDev &dm, int line,
                                             int pos, float L) {
    int idx;
    if (sg.dir == 0) idx = sg.field == 0 ? dm.off_vx + line * (dm.nx + 1) + pos : dm.off_px + line * dm.nx + pos;
    else idx = sg.field == 0 ? dm.off_vy + pos * dm.nx + line : dm.off_py + pos * dm.nx + line;
    float upd = sg.field == 0 ? L / dm.rho : L * dm.rho * a.c2;
    float v = a.tn[idx] - a.s * upd;
    a.out[idx] = v;
    if (a.last) a.tn[idx] = dm.is_pml ? v * a.pml[idx] : v;
}

template <int F, int LAYOUT>
__device__ __forceinline__ void derivative_tile(const StageArgs &a, const TileDev &tile, float2 *buf0, float2 *buf1) {
    constexpr int T = F / 8;
    constexpr int NT = F >= 4096 ? F / 8 : kSmallThreads;
    constexpr int P = NT / T;  // line pairs per block
    const int tid = threadIdx.x;
    const int p = LAYOUT == 0 ? tid / T : tid % P;
    const int t = LAYOUT == 0 ? tid % T : tid / P;
    const SegDev sg = a.segs[tile.seg];
    const DomainDev dm = a.doms[sg.domain];
    const int l0 = tile.line0 + 2 * p, l1 = l0 + 1;
    const bool ok0 = 2 * p < tile.count, ok1 = 2 * p + 1 < tile.count;
    float2 v[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int i = t + m * T;
        v[m].x = ok0 ? line_value(a, sg, dm, l0, i) : 0.f;
        v[m].y = ok1 ? line_value(a, sg, dm, l1, i) : 0.f;
    }
    const float2 *tw = a.ctab + sg.tw_off;
    fft_run<F, P, LAYOUT, false>(v, p, t, buf0, buf1, tw);
    const float2 *H = a.ctab + sg.h_off;
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = cmul(v[m], H[t + m * T]);
    __syncthreads();
    fft_run<F, P, LAYOUT, true>(v, p, t, buf0, buf1, tw);
    const int w = sg.wlen;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int i = t + m * T;
        if (i >= w && i < w + sg.n_out) {
            if (ok0) write_update(a, sg, dm, l0, i - w, v[m].x);
            if (ok1) write_update(a, sg, dm, l1, i - w, v[m].y);
        }
    }
}

__device__ void passthrough_tile(const StageArgs &a, const TileDev &tile) {
    const DomainDev dm = a.doms[tile.domain];
    const int base = field_offset(dm, tile.field);
    for (int e = threadIdx.x; e < tile.count; e += blockDim.x) {
        const int idx = base + tile.line0 + e;
        float v = a.tn[idx];
        a.out[idx] = v;
        if (a.last && dm.is_pml) a.tn[idx] = v * a.pml[idx];
    }
}

template <int F>
__device__ __forceinline__ void run_tile_F(const StageArgs &a, const TileDev &tile, float2 *lds) {
    constexpr int NT = F >= 4096 ? F / 8 : kSmallThreads;
    constexpr int P = NT / (F / 8);
    const SegDev &sg = a.segs[tile.seg];
    if (sg.dir == 0) {
        using Lm = LdsMap<F, P, 0>;
        derivative_tile<F, 0>(a, tile, lds, lds + Lm::SIZE);
    } else {
        using Lm = LdsMap<F, P, 1>;
        derivative_tile<F, 1>(a, tile, lds, lds + Lm::SIZE);
    }
}

// two LDS buffers of P * F = 8 * kSmallThreads complex values
constexpr int lds_small_floats2() { return 2 * 8 * kSmallThreads; }

__global__ void __launch_bounds__(kSmallThreads) stage_small_kernel(StageArgs a) {
    __shared__ float2 lds[lds_small_floats2()];
    const TileDev tile = a.tiles[blockIdx.x];
    if (tile.seg < 0) {
        passthrough_tile(a, tile);
        return;
    }
    switch (a.segs[tile.seg].F) {
        case 128: run_tile_F<128>(a, tile, lds); break;
        case 256: run_tile_F<256>(a, tile, lds); break;
        case 512: run_tile_F<512>(a, tile, lds); break;
        case 1024: run_tile_F<1024>(a, tile, lds); break;
        case 2048: run_tile_F<2048>(a, tile, lds); break;
        default: break;
    }
}

template <int F>
__global__ void __launch_bounds__(F / 8) stage_big_kernel(StageArgs a) {
    extern __shared__ float2 lds_dyn[];
    const TileDev tile = a.tiles[blockIdx.x];
    run_tile_F<F>(a, tile, lds_dyn);
}

__global__ void receiver_kernel(const float *__restrict__ arena, const int2 *__restrict__ cells, int nrec,
                                float *__restrict__ samples, int slot) {
    // cells[r] = (px slot index, py slot index); p0 = px0 + py0 (kernel/Solver.cpp:114-116)
    for (int r = threadIdx.x; r < nrec; r += blockDim.x) {
        int2 c = cells[r];
        samples[(size_t)slot * nrec + r] = arena[c.x] + arena[c.y];
    }
}

__global__ void p0_materialise_kernel(const float *__restrict__ src, float *__restrict__ p0i, const int4 *ranges,
                                      int nr) {
    // ranges[k] = (px offset, py offset, count, -)
    for (int k = blockIdx.y; k < nr; k += gridDim.y) {
        int4 r = ranges[k];
        for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < r.z; e += gridDim.x * blockDim.x)
            p0i[r.x + e] = src[r.x + e] + src[r.y + e];
    }
}

// ------------------------------------------------------------- host side

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(_e);               \
            return PSTD_ERR_DEVICE;                                                     \
        }                                                                               \
    } while (0)

int next_pow2(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

}  // namespace

struct pstd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool planned = false;
    // plan
    std::vector<DomainDev> doms;
    std::vector<SegDev> segs;
    std::vector<pstd_plan_domain> plan_domains;
    std::vector<TileDev> tiles_small, tiles_4096, tiles_8192;
    int arena_floats = 0;
    float dt = 0, c2 = 0;
    float alpha[6] = {0};
    int nrec = 0;
    // device buffers
    float *arena = nullptr;  // 4 arenas back to back: TN, S0, S1, P0I
    float *pml = nullptr;
    float2 *ctab = nullptr;
    float *ftab = nullptr;
    DomainDev *d_doms = nullptr;
    SegDev *d_segs = nullptr;
    TileDev *d_tiles = nullptr;  // small | 4096 | 8192
    int2 *d_cells = nullptr;
    int4 *d_ranges = nullptr;
    int n_ranges = 0;
    float *d_samples = nullptr;
    int sample_cap = 0;
    int steps_done = 0;
    bool p0_explicit = true;

    float *TN() const { return arena; }
    float *S(int i) const { return arena + (size_t)(1 + i) * arena_floats; }
    float *P0I() const { return arena + (size_t)3 * arena_floats; }

    void release() {
        (void)hipFree(arena);
        (void)hipFree(pml);
        (void)hipFree(ctab);
        (void)hipFree(ftab);
        (void)hipFree(d_doms);
        (void)hipFree(d_segs);
        (void)hipFree(d_tiles);
        (void)hipFree(d_cells);
        (void)hipFree(d_ranges);
        (void)hipFree(d_samples);
        arena = pml = ftab = d_samples = nullptr;
        ctab = nullptr;
        d_doms = nullptr;
        d_segs = nullptr;
        d_tiles = nullptr;
        d_cells = nullptr;
        d_ranges = nullptr;
        planned = false;
    }
};

namespace {

// Derivative filter H[k] of a line of FFT length F (device plan cache).
// derfact of WisdomCache::discretize_wave_numbers (kernel/core/WisdomCache.cpp:53-75)
// for a table of length L: k_m = m * (pi/dx) / (L/2); pressure
// exp(-i k dx/2) i k, velocity exp(+i k dx/2) i k.  spatderp3 multiplies bins
// 0..F/2 (kernel/core/kernel_functions.cpp:234, 294) and FFTW's c2r ignores the
// imaginary parts of bins 0 and F/2; extending Hermitian-symmetrically lets one
// complex transform carry two real lines.  1/F (kernel_functions.cpp:243, 306)
// is folded in.  Assembled in double.
void build_filter(int F, int L, int field, double dx, std::vector<float2> &out) {
    out.assign(F, make_float2(0, 0));
    const double kmax = M_PI / dx;
    const double dk = kmax / (L / 2);
    std::vector<std::complex<double>> d(F / 2 + 1);
    for (int m = 0; m <= F / 2; ++m) {
        double k = m * dk;
        std::complex<double> ik(0, k);
        std::complex<double> sh = std::exp(std::complex<double>(0, (field == 0 ? -1 : 1) * k * dx * 0.5));
        d[m] = sh * ik;
    }
    d[0] = 0;
    d[F / 2] = std::complex<double>(d[F / 2].real(), 0);
    for (int m = 0; m <= F / 2; ++m) {
        std::complex<double> h = d[m] / (double)F;
        out[m] = make_float2((float)h.real(), (float)h.imag());
        if (m > 0 && m < F / 2) out[F - m] = make_float2((float)h.real(), (float)-h.imag());
    }
}

void build_twiddles(int F, std::vector<float2> &out) {
    out.resize(F);
    for (int i = 0; i < F; ++i) {
        double a = -2.0 * M_PI * (double)i / (double)F;
        out[i] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
}

// get_window_coefficients (kernel/core/kernel_functions.cpp:327-333) in double.
void build_window(int w, int patch_error, std::vector<float> &out) {
    out.resize(2 * w + 1);
    double alpha = (patch_error - 40) / 20.0 + 1;
    for (int i = 0; i <= 2 * w; ++i) {
        double x = (double)(i - w) / w;
        out[i] = (float)std::exp(-std::log(10.0) * alpha * std::pow(x, 6));
    }
}

int launch_stage(pstd_ctx *ctx, int k, bool first_of_step) {
    StageArgs a{};
    a.doms = ctx->d_doms;
    a.segs = ctx->d_segs;
    a.ctab = ctx->ctab;
    a.ftab = ctx->ftab;
    const int in = (k + 1) & 1, outi = k & 1;
    if (k == 0) {
        if (ctx->p0_explicit) {
            a.p0a = ctx->P0I();
            a.p0b = nullptr;
        } else {
            a.p0a = ctx->S(1);
            a.p0b = ctx->S(1);
        }
        a.vsrc = ctx->TN();
    } else {
        a.p0a = ctx->S(in);
        a.p0b = ctx->S(in);
        a.vsrc = ctx->S(in);
    }
    (void)first_of_step;
    a.tn = ctx->TN();
    a.out = ctx->S(outi);
    a.pml = ctx->pml;
    a.s = ctx->dt * ctx->alpha[k];
    a.c2 = ctx->c2;
    a.last = k == 5;
    const int ns = (int)ctx->tiles_small.size(), n4 = (int)ctx->tiles_4096.size(), n8 = (int)ctx->tiles_8192.size();
    if (ns) {
        a.tiles = ctx->d_tiles;
        a.ntiles = ns;
        hipLaunchKernelGGL(stage_small_kernel, dim3(ns), dim3(kSmallThreads), 0, ctx->stream, a);
    }
    if (n4) {
        a.tiles = ctx->d_tiles + ns;
        a.ntiles = n4;
        size_t lds = 2 * sizeof(float2) * LdsMap<4096, 1, 0>::SIZE;
        hipLaunchKernelGGL(stage_big_kernel<4096>, dim3(n4), dim3(512), lds, ctx->stream, a);
    }
    if (n8) {
        a.tiles = ctx->d_tiles + ns + n4;
        a.ntiles = n8;
        size_t lds = 2 * sizeof(float2) * LdsMap<8192, 1, 0>::SIZE;
        hipLaunchKernelGGL(stage_big_kernel<8192>, dim3(n8), dim3(1024), lds, ctx->stream, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->err = std::string("stage launch: ") + hipGetErrorString(e);
        return PSTD_ERR_DEVICE;
    }
    return PSTD_OK;
}

int field_size(const pstd_plan_domain &d, int f) {
    switch (f) {
        case PSTD_FIELD_VX: return d.ny * (d.nx + 1);
        case PSTD_FIELD_VY: return (d.ny + 1) * d.nx;
        default: return d.nx * d.ny;
    }
}

int field_off(const DomainDev &d, int f) {
    switch (f) {
        case PSTD_FIELD_VX: return d.off_vx;
        case PSTD_FIELD_VY: return d.off_vy;
        case PSTD_FIELD_PX: return d.off_px;
        case PSTD_FIELD_PY: return d.off_py;
        default: return d.off_px;  // p0 lives in the px slot of P0I
    }
}

int materialise_p0(pstd_ctx *ctx) {
    // P0I <- S1.px + S1.py for every domain (the current, stale p0)
    if (ctx->p0_explicit) return PSTD_OK;
    if (ctx->steps_done == 0) return PSTD_OK;
    hipLaunchKernelGGL(p0_materialise_kernel, dim3(64, std::min(ctx->n_ranges, 64)), dim3(256), 0, ctx->stream,
                       ctx->S(1), ctx->P0I(), ctx->d_ranges, ctx->n_ranges);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        ctx->err = std::string("p0 launch: ") + hipGetErrorString(e);
        return PSTD_ERR_DEVICE;
    }
    ctx->p0_explicit = true;
    return PSTD_OK;
}

}  // namespace

extern "C" {

int pstd_ctx_create(int device, pstd_ctx **out) {
    if (!out) return PSTD_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return PSTD_ERR_DEVICE;
    if (device < 0 || device >= n) return PSTD_ERR_INVALID;
    pstd_ctx *ctx = new pstd_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return PSTD_ERR_DEVICE;
    }
    // the F = 8192 class needs 128 KiB of LDS per workgroup
    if (hipFuncSetAttribute((const void *)stage_big_kernel<8192>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            2 * sizeof(float2) * 8192) != hipSuccess ||
        hipFuncSetAttribute((const void *)stage_big_kernel<4096>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            2 * sizeof(float2) * 4096) != hipSuccess) {
        ctx->err = "cannot reserve LDS for the long-line kernels";
    }
    *out = ctx;
    return PSTD_OK;
}

void pstd_ctx_destroy(pstd_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->release();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *pstd_ctx_last_error(const pstd_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pstd_ctx_upload_plan(pstd_ctx *ctx, const pstd_plan_desc *plan) {
    if (!ctx || !plan) return PSTD_ERR_INVALID;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->release();
    ctx->doms.clear();
    ctx->segs.clear();
    ctx->tiles_small.clear();
    ctx->tiles_4096.clear();
    ctx->tiles_8192.clear();
    ctx->plan_domains.assign(plan->domains, plan->domains + plan->n_domains);
    // arena layout
    int off = 0;
    auto align = [](int v) { return (v + 63) & ~63; };
    for (int i = 0; i < plan->n_domains; ++i) {
        const pstd_plan_domain &p = plan->domains[i];
        if (p.nx <= 0 || p.ny <= 0) {
            ctx->err = "domain with empty extent";
            return PSTD_ERR_INVALID;
        }
        DomainDev d{};
        d.nx = p.nx;
        d.ny = p.ny;
        d.rho = p.rho;
        d.is_pml = p.is_pml;
        d.off_vx = off;
        off = align(off + p.ny * (p.nx + 1));
        d.off_vy = off;
        off = align(off + (p.ny + 1) * p.nx);
        d.off_px = off;
        off = align(off + p.ny * p.nx);
        d.off_py = off;
        off = align(off + p.ny * p.nx);
        ctx->doms.push_back(d);
    }
    ctx->arena_floats = std::max(off, 64);
    ctx->dt = plan->dt;
    ctx->c2 = plan->c1 * plan->c1;
    for (int k = 0; k < 6; ++k) ctx->alpha[k] = plan->rk_alpha[k];
    // device plan cache: filters, twiddles, windows
    std::vector<float2> ctab;
    std::vector<float> ftab;
    std::map<std::tuple<int, int, int>, int> filt_cache;
    std::map<int, int> tw_cache;
    std::map<int, int> win_cache;
    const double dx = plan->dx;
    for (int si = 0; si < plan->n_segments; ++si) {
        const pstd_plan_segment &s = plan->segments[si];
        if (s.domain < 0 || s.domain >= plan->n_domains || s.side1 >= plan->n_domains || s.side2 >= plan->n_domains ||
            (s.direction != 0 && s.direction != 1) || (s.field != 0 && s.field != 1) || s.wlen <= 0) {
            ctx->err = "invalid segment " + std::to_string(si);
            return PSTD_ERR_INVALID;
        }
        const DomainDev &dm = ctx->doms[s.domain];
        const int n = s.direction == 0 ? dm.nx : dm.ny;
        const int nlines = s.direction == 0 ? dm.ny : dm.nx;
        if (s.line_begin < 0 || s.line_end > nlines || s.line_begin >= s.line_end) {
            ctx->err = "segment line range outside its domain";
            return PSTD_ERR_INVALID;
        }
        SegDev g{};
        g.domain = s.domain;
        g.dir = s.direction;
        g.field = s.field;
        g.side1 = s.side1;
        g.side2 = s.side2;
        g.off1 = s.side1_offset;
        g.off2 = s.side2_offset;
        g.wlen = s.wlen;
        g.n_main = s.field == 0 ? n : n + 1;
        g.n_out = s.field == 0 ? n + 1 : n;
        g.F = next_pow2(g.n_main + 2 * s.wlen);
        g.c_s1 = s.c_side1;
        g.c_m1 = s.c_main1;
        g.c_s2 = s.c_side2;
        g.c_m2 = s.c_main2;
        if (g.F < 128 || g.F > 8192) {
            ctx->err = "FFT length " + std::to_string(g.F) + " outside 128..8192";
            return PSTD_ERR_UNSUPPORTED;
        }
        if (s.table_len < g.F || (s.table_len & (s.table_len - 1))) {
            ctx->err = "wavenumber table shorter than the FFT length";
            return PSTD_ERR_INVALID;
        }
        if (s.wlen > n) {
            ctx->err = "window longer than the domain";
            return PSTD_ERR_INVALID;
        }
        // neighbour bounds: lines and halo columns must exist in the neighbour
        for (int side = 0; side < 2; ++side) {
            int nb = side == 0 ? s.side1 : s.side2;
            int o = side == 0 ? s.side1_offset : s.side2_offset;
            if (nb < 0) continue;
            const DomainDev &dn = ctx->doms[nb];
            int nlen = (s.direction == 0 ? dn.nx : dn.ny) + (s.field == 1 ? 1 : 0);
            int nlin = s.direction == 0 ? dn.ny : dn.nx;
            if (s.line_begin + o < 0 || s.line_end + o > nlin || nlen < s.wlen + (s.field == 1 ? 1 : 0)) {
                ctx->err = "segment reads outside its neighbour domain";
                return PSTD_ERR_INVALID;
            }
        }
        auto fk = std::make_tuple(g.F, s.table_len, s.field);
        auto it = filt_cache.find(fk);
        if (it == filt_cache.end()) {
            std::vector<float2> h;
            build_filter(g.F, s.table_len, s.field, dx, h);
            it = filt_cache.emplace(fk, (int)ctab.size()).first;
            ctab.insert(ctab.end(), h.begin(), h.end());
        }
        g.h_off = it->second;
        auto tt = tw_cache.find(g.F);
        if (tt == tw_cache.end()) {
            std::vector<float2> tw;
            build_twiddles(g.F, tw);
            tt = tw_cache.emplace(g.F, (int)ctab.size()).first;
            ctab.insert(ctab.end(), tw.begin(), tw.end());
        }
        g.tw_off = tt->second;
        auto wt = win_cache.find(s.wlen);
        if (wt == win_cache.end()) {
            std::vector<float> w;
            build_window(s.wlen, (int)plan->patch_error, w);
            wt = win_cache.emplace(s.wlen, (int)ftab.size()).first;
            ftab.insert(ftab.end(), w.begin(), w.end());
        }
        g.win_off = wt->second;
        ctx->segs.push_back(g);
        // tiles: 2 lines per pair, P pairs per block
        const int P = g.F >= 4096 ? 1 : kSmallThreads / (g.F / 8);
        const int per = 2 * P;
        for (int l = s.line_begin; l < s.line_end; l += per) {
            TileDev t{};
            t.seg = si;
            t.line0 = l;
            t.count = std::min(per, s.line_end - l);
            if (g.F == 4096) ctx->tiles_4096.push_back(t);
            else if (g.F == 8192) ctx->tiles_8192.push_back(t);
            else ctx->tiles_small.push_back(t);
        }
    }
    // passthrough tiles for fields no segment updates (rigid domains, directions
    // openPSTD does not update)
    std::vector<std::vector<char>> covered(plan->n_domains, std::vector<char>(4, 0));
    for (const SegDev &g : ctx->segs) {
        int f = g.dir == 0 ? (g.field == 0 ? PSTD_FIELD_VX : PSTD_FIELD_PX) : (g.field == 0 ? PSTD_FIELD_VY : PSTD_FIELD_PY);
        covered[g.domain][f - 1] = 1;
    }
    for (int i = 0; i < plan->n_domains; ++i)
        for (int f = PSTD_FIELD_PX; f <= PSTD_FIELD_VY; ++f) {
            if (covered[i][f - 1]) continue;
            int n = field_size(plan->domains[i], f);
            for (int e = 0; e < n; e += kPassChunk) {
                TileDev t{};
                t.seg = -1;
                t.line0 = e;
                t.count = std::min(kPassChunk, n - e);
                t.domain = i;
                t.field = f;
                ctx->tiles_small.push_back(t);
            }
        }
    if (ctab.empty()) ctab.push_back(make_float2(0, 0));
    if (ftab.empty()) ftab.push_back(0);
    // PML factor arena (1 where no PML)
    std::vector<float> pml((size_t)ctx->arena_floats, 1.f);
    for (int i = 0; i < plan->n_domains; ++i) {
        const pstd_plan_domain &p = plan->domains[i];
        if (!p.is_pml) continue;
        const DomainDev &d = ctx->doms[i];
        const float *src[4] = {p.pml_vx, p.pml_vy, p.pml_px, p.pml_py};
        const int offs[4] = {d.off_vx, d.off_vy, d.off_px, d.off_py};
        const int sizes[4] = {field_size(p, PSTD_FIELD_VX), field_size(p, PSTD_FIELD_VY), field_size(p, PSTD_FIELD_PX),
                              field_size(p, PSTD_FIELD_PY)};
        for (int f = 0; f < 4; ++f) {
            if (!src[f]) {
                ctx->err = "PML domain without factor arrays";
                return PSTD_ERR_INVALID;
            }
            std::memcpy(pml.data() + offs[f], src[f], sizeof(float) * sizes[f]);
        }
    }
    // receivers
    std::vector<int2> cells;
    for (int r = 0; r < plan->n_receivers; ++r) {
        const pstd_plan_receiver &rc = plan->receivers[r];
        if (rc.domain < 0 || rc.domain >= plan->n_domains || rc.index < 0 ||
            rc.index >= plan->domains[rc.domain].nx * plan->domains[rc.domain].ny) {
            ctx->err = "receiver outside its domain";
            return PSTD_ERR_INVALID;
        }
        const DomainDev &d = ctx->doms[rc.domain];
        cells.push_back(make_int2(d.off_px + rc.index, d.off_py + rc.index));
    }
    ctx->nrec = (int)cells.size();
    std::vector<int4> ranges;
    for (const DomainDev &d : ctx->doms) ranges.push_back(make_int4(d.off_px, d.off_py, d.nx * d.ny, 0));
    ctx->n_ranges = (int)ranges.size();
    // allocate + upload
    const size_t af = (size_t)ctx->arena_floats;
    HIP_TRY(hipMalloc(&ctx->arena, 4 * af * sizeof(float)));
    HIP_TRY(hipMemsetAsync(ctx->arena, 0, 4 * af * sizeof(float), ctx->stream));
    HIP_TRY(hipMalloc(&ctx->pml, af * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(ctx->pml, pml.data(), af * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMalloc(&ctx->ctab, ctab.size() * sizeof(float2)));
    HIP_TRY(hipMemcpyAsync(ctx->ctab, ctab.data(), ctab.size() * sizeof(float2), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMalloc(&ctx->ftab, ftab.size() * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(ctx->ftab, ftab.data(), ftab.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMalloc(&ctx->d_doms, std::max<size_t>(1, ctx->doms.size()) * sizeof(DomainDev)));
    if (!ctx->doms.empty())
        HIP_TRY(hipMemcpyAsync(ctx->d_doms, ctx->doms.data(), ctx->doms.size() * sizeof(DomainDev),
                               hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMalloc(&ctx->d_segs, std::max<size_t>(1, ctx->segs.size()) * sizeof(SegDev)));
    if (!ctx->segs.empty())
        HIP_TRY(hipMemcpyAsync(ctx->d_segs, ctx->segs.data(), ctx->segs.size() * sizeof(SegDev),
                               hipMemcpyHostToDevice, ctx->stream));
    std::vector<TileDev> all = ctx->tiles_small;
    all.insert(all.end(), ctx->tiles_4096.begin(), ctx->tiles_4096.end());
    all.insert(all.end(), ctx->tiles_8192.begin(), ctx->tiles_8192.end());
    HIP_TRY(hipMalloc(&ctx->d_tiles, std::max<size_t>(1, all.size()) * sizeof(TileDev)));
    if (!all.empty())
        HIP_TRY(hipMemcpyAsync(ctx->d_tiles, all.data(), all.size() * sizeof(TileDev), hipMemcpyHostToDevice,
                               ctx->stream));
    HIP_TRY(hipMalloc(&ctx->d_cells, std::max<size_t>(1, cells.size()) * sizeof(int2)));
    if (!cells.empty())
        HIP_TRY(hipMemcpyAsync(ctx->d_cells, cells.data(), cells.size() * sizeof(int2), hipMemcpyHostToDevice,
                               ctx->stream));
    HIP_TRY(hipMalloc(&ctx->d_ranges, std::max<size_t>(1, ranges.size()) * sizeof(int4)));
    if (!ranges.empty())
        HIP_TRY(hipMemcpyAsync(ctx->d_ranges, ranges.data(), ranges.size() * sizeof(int4), hipMemcpyHostToDevice,
                               ctx->stream));
    ctx->sample_cap = 4096;
    HIP_TRY(hipMalloc(&ctx->d_samples, (size_t)ctx->sample_cap * std::max(1, ctx->nrec) * sizeof(float)));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->steps_done = 0;
    ctx->p0_explicit = true;
    ctx->planned = true;
    return PSTD_OK;
}

int pstd_ctx_set_field(pstd_ctx *ctx, int domain, int field, const float *host, size_t count) {
    if (!ctx || !ctx->planned || !host) return PSTD_ERR_INVALID;
    if (domain < 0 || domain >= (int)ctx->doms.size() || field < 0 || field > 4) return PSTD_ERR_INVALID;
    const pstd_plan_domain &p = ctx->plan_domains[domain];
    if (count != (size_t)field_size(p, field)) {
        ctx->err = "field size mismatch";
        return PSTD_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    float *dst;
    if (field == PSTD_FIELD_P0) {
        int rc = materialise_p0(ctx);
        if (rc) return rc;
        dst = ctx->P0I() + ctx->doms[domain].off_px;
    } else {
        dst = ctx->TN() + field_off(ctx->doms[domain], field);
    }
    HIP_TRY(hipMemcpyAsync(dst, host, count * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PSTD_OK;
}

int pstd_ctx_get_field(pstd_ctx *ctx, int domain, int field, float *host, size_t count) {
    if (!ctx || !ctx->planned || !host) return PSTD_ERR_INVALID;
    if (domain < 0 || domain >= (int)ctx->doms.size() || field < 0 || field > 4) return PSTD_ERR_INVALID;
    const pstd_plan_domain &p = ctx->plan_domains[domain];
    if (count != (size_t)field_size(p, field)) {
        ctx->err = "field size mismatch";
        return PSTD_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    const DomainDev &d = ctx->doms[domain];
    if (field == PSTD_FIELD_P0) {
        if (ctx->p0_explicit) {
            HIP_TRY(hipMemcpyAsync(host, ctx->P0I() + d.off_px, count * sizeof(float), hipMemcpyDeviceToHost,
                                   ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
        } else {
            std::vector<float> py(count);
            HIP_TRY(hipMemcpyAsync(host, ctx->S(1) + d.off_px, count * sizeof(float), hipMemcpyDeviceToHost,
                                   ctx->stream));
            HIP_TRY(hipMemcpyAsync(py.data(), ctx->S(1) + d.off_py, count * sizeof(float), hipMemcpyDeviceToHost,
                                   ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            for (size_t i = 0; i < count; ++i) host[i] = host[i] + py[i];
        }
        return PSTD_OK;
    }
    HIP_TRY(hipMemcpyAsync(host, ctx->TN() + field_off(d, field), count * sizeof(float), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PSTD_OK;
}

int pstd_ctx_read_frame(pstd_ctx *ctx, int domain, float *out, size_t count) {
    return pstd_ctx_get_field(ctx, domain, PSTD_FIELD_P0, out, count);
}

static int advance_impl(pstd_ctx *ctx, int steps, float *samples) {
    if (!ctx || !ctx->planned) return PSTD_ERR_INVALID;
    if (steps < 0) return PSTD_ERR_INVALID;
    HIP_TRY(hipSetDevice(ctx->device));
    int done = 0;
    while (done < steps) {
        const int chunk = std::min(steps - done, ctx->sample_cap);
        for (int s = 0; s < chunk; ++s) {
            for (int k = 0; k < 6; ++k) {
                int rc = launch_stage(ctx, k, s == 0);
                if (rc) return rc;
                if (k == 0) ctx->p0_explicit = false;
            }
            if (ctx->nrec > 0) {
                hipLaunchKernelGGL(receiver_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->S(1), ctx->d_cells,
                                   ctx->nrec, ctx->d_samples, s);
                HIP_TRY(hipGetLastError());
            }
            ctx->steps_done++;
        }
        if (samples && ctx->nrec > 0) {
            HIP_TRY(hipMemcpyAsync(samples + (size_t)done * ctx->nrec, ctx->d_samples,
                                   (size_t)chunk * ctx->nrec * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
        }
        done += chunk;
    }
    return PSTD_OK;
}

int pstd_ctx_advance(pstd_ctx *ctx, int steps, float *samples) {
    int rc = advance_impl(ctx, steps, samples);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PSTD_OK;
}

int pstd_ctx_advance_async(pstd_ctx *ctx, int steps) { return advance_impl(ctx, steps, nullptr); }

int pstd_ctx_synchronize(pstd_ctx *ctx) {
    if (!ctx) return PSTD_ERR_INVALID;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PSTD_OK;
}

int pstd_ctx_steps_done(const pstd_ctx *ctx) { return ctx ? ctx->steps_done : -1; }

}  // extern "C"
