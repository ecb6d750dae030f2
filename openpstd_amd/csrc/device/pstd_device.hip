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
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../../include/openpstd_amd.h"
#include "../host/plan_tables.h"
#include "fft_line.h"

using namespace pstdamd::dev;

// Scratch of the allocation-free advance (pstd_ctx_advance_static): the
// GEMM form's per-scene maxima, 3 slots x 2 classes x kStaticScenes scenes
// (24 KiB), one copy per device (each device has its own address, resolved
// with hipGetSymbolAddress), shared by every call on that device.
constexpr int kStaticScenes = 1024;
static __device__ __align__(256) unsigned scratch[3 * 2 * kStaticScenes];

namespace {

constexpr int kPassChunk = 2048;    // elements per passthrough tile
constexpr int kMaxDomainFloats = 1 << 29;  // one domain block (4 fields) per buffer resource

// A domain's four fields (vx, vy, px, py, row-major) form one block at the same
// offset of every arena (TN, S0, S1, P0I and the PML factors).
struct DomainDev {
    int nx, ny;
    long long base;                 // element offset of the block in an arena
    int size;                       // elements of the block
    int f_vx, f_vy, f_px, f_py;     // field offsets within the block
    float rho;
    int is_pml;
};

// One work item of a stage launch; everything a workgroup needs in one
// scalar-loadable record.
//  kind 0: derivative tile — `count` lines of a segment (Domain::calc,
//          kernel/core/Domain.cpp:149-370): gather, derivative, update.
//  kind 1: passthrough — elements [line0, line0+count) of one field that no
//          segment updates (rigid domains, directions openPSTD skips).
//  kind 2: receiver sampling (first stage of a step).
//  kind 3: the segment addressing of a GEMM-form tile (GemmTile::t).
// Memory is addressed through buffer resources over one domain block (base
// `*_dom`, `*_size` elements): value (line, pos) of a source field is element
// a + line*ls + pos*ps of the block (+ element b of the second array for
// p0 = px0 + py0).  Offsets past the block read 0 and drop stores, which is how
// absent neighbours, the zero padding and partial tiles are expressed.
struct TileDesc {
    int kind;
    int F;
    int line0, count;
    int wlen, wlog, n_main, n_out, vel, dir;  // wlog: log2 of the halo slot (2^wlog >= wlen)
    int h_off, tw_off, win_off;
    int tw2_off;  // F = 8192: the 4096-point twiddle table (its two halves run as 4096-point transforms)
    long long m_dom, s1_dom, s2_dom;
    int m_size, s1_size, s2_size;
    int m_a, m_b, m_ls, m_ps;
    int s1_a, s1_b, s1_ls, s1_ps, s1_line, s1_len;  // s1_len < 0: no neighbour (zeros)
    int s2_a, s2_b, s2_ls, s2_ps, s2_line, s2_len;
    float c_s1, c_m1, c_s2, c_m2;
    int o_f, o_ls, o_ps;  // output field (vx, px, vy or py) within the main block
    float coef;           // 1/rho (velocity update) or rho*c1^2 (pressure update)
    int is_pml;
    int scene;            // scene of the batch (per-scene maxima of the GEMM form's scaling)
};

struct ReceiverCell {
    long long px, py;  // arena element offsets of the cell's px0 and py0; px < 0: spectral receiver
};

struct StageArgs {
    const TileDesc *__restrict__ tiles;
    const float2 *ctab;
    const float *ftab;
    const float *p0a;  // p0 = p0a[...] + (p0b ? p0b[...] : 0)
    const float *p0b;
    const float *vsrc; // current vx / vy
    float *tn;
    float *out;
    const float *pml;  // PML factor arena (same layout); 1 for non-PML domains
    float s;           // dt * alpha_k
    int last;          // last RK stage: also write tn = out * pml
    // receiver sampling of the previous step (first stage of a step): p0 = px + py
    // of the previous step's raw stage-5 output (kernel/Solver.cpp:86-91)
    const ReceiverCell *cells;
    int nrec;
    float *samples;    // samples[slot * nrec + r], slot < 0: no sampling in this launch
    int slot;
    int remap;         // XCD-contiguous tile order (tile_index)
    // GEMM form (umax == nullptr otherwise): per-scene maxima of the written
    // values in three rotating slots (write slot_w, read slot_r, clear slot_z)
    unsigned *umax;
    int slot_w, slot_r, slot_z, n_scenes;
    const void *dmat;  // derivative matrices (2-byte elements)
};

// ------------------------------------------------------------- device code

typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int kOOB = (int)0x80000000u;  // byte offset past every block: loads 0, drops stores

__device__ __forceinline__ Rsrc block_rsrc(const float *arena, long long dom, int size) {
    // raw buffer (stride 0) over `size` floats; word 3 as for every gfx9 raw buffer
    return __builtin_amdgcn_make_buffer_rsrc((void *)(arena + dom), (short)0, size * 4, 0x00020000);
}
__device__ __forceinline__ float bload(Rsrc r, int off_bytes) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off_bytes, 0, 0));
}
__device__ __forceinline__ void bstore(Rsrc r, int off_bytes, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off_bytes, 0, 0);
}
// two adjacent floats (4-byte aligned suffices for buffer accesses)
__device__ __forceinline__ float2 bload2(Rsrc r, int off_bytes) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off_bytes, 0, 0));
}
__device__ __forceinline__ void bstore2(Rsrc r, int off_bytes, float x, float y) {
    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, make_float2(x, y)), r, off_bytes, 0, 0);
}

// Geometry of a derivative tile: F-point lines (run time), NT threads,
// T = F/8 threads per line pair, P = NT/T pairs; LAYOUT (compile time) 0: X
// lines (rows; lanes walk positions), 1: Y lines (columns; lanes walk the
// pairs).  The gather and update code is shared by all F (one copy per
// layout keeps the kernel small enough for the instruction cache); only the
// transforms are specialised per F.
template <int LAYOUT>
struct TileGeo {
    int F, T, P, logT, logP;
    __device__ TileGeo(int F_, int NT) : F(F_), T(F_ >> 3), P(NT / (F_ >> 3)) {
        logT = __ffs(T) - 1;
        logP = __ffs(P) - 1;
    }
    __device__ int pair(int tid) const { return LAYOUT == 0 ? tid >> logT : tid & (P - 1); }
    __device__ int lane_t(int tid) const { return LAYOUT == 0 ? tid & (T - 1) : tid >> logP; }
    // raw staging of the 2P lines in LDS (float): line-major for rows,
    // position-major for columns, so that the wavefront's lanes stay contiguous.
    // Rows use a pitch of 9F/8 floats: the two lines of pair p then fill exactly
    // the slots pair p's transform uses in the second buffer (its own, so the
    // transforms of rows need no workgroup barrier after the staging)
    __device__ int line_base(int ll) const { return LAYOUT == 0 ? (ll << (logT + 3)) + (ll << logT) : ll; }
    __device__ int pos_step() const { return LAYOUT == 0 ? 1 : 2 * P; }
    __device__ int stage_at(int ll, int pos) const { return line_base(ll) + pos * pos_step(); }
};

// The windowed lines of one tile (spatderp3, kernel/core/kernel_functions.cpp:
// 210-216 pressure, :263-269 velocity):  [dom1 | main | dom3 | 0], where
// dom1[j] = (side1[tail + j] * c_s1 + main[reflected] * c_m1) * window[j] and
// dom3 likewise.  Thread (p, t) loads the main values it will transform
// (positions t + m*T of lines 2p and 2p+1); the 2 x w halo values of every
// line are loaded by the tile's threads in turn; both go to LDS so that halo
// positions can combine them with the reflected main values.  All loads are
// issued before the first wait.
struct GatherRegs {
    float ra[8][2], rb[8][2];
    float h1a[4], h1b[4], h2a[4], h2b[4];
    int hpos1[4], hpos2[4];
    float win;  // window coefficient threadIdx.x (threads 0..2w)
};

template <int NT, int LAYOUT>
__device__ __forceinline__ void gather_issue(const StageArgs &a, const TileDesc &td, const TileGeo<LAYOUT> &g, int p,
                                             int t, GatherRegs &R) {
    const int w = td.wlen, nm = td.n_main, vel = td.vel, count = td.count;
    __builtin_assume(w > 0 && w <= 32);
    // the window, with the lines (not after them: one memory latency less)
    R.win = (int)threadIdx.x <= 2 * w ? a.ftab[td.win_off + threadIdx.x] : 0.f;
    const float *A = vel ? a.vsrc : a.p0a;
    const bool hasB = !vel && a.p0b != nullptr;
    const float *B = hasB ? a.p0b : A;
    // resources: main, side1, side2 blocks; B shifted so that A's offsets address it
    const Rsrc rmA = block_rsrc(A, td.m_dom, td.m_size);
    const Rsrc rmB = block_rsrc(B, td.m_dom + (td.m_b - td.m_a), td.m_size - (td.m_b - td.m_a));
    const int la = td.line0 + 2 * p;
    const int pstep = td.m_ps * 4;
    const int lbase = (td.m_a + la * td.m_ls) * 4 + (t - w) * pstep, lstep = td.m_ls * 4, mstep = g.T * pstep;
    int off[8][2];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int e = t + m * g.T - w;
        const bool ok = e >= 0 && e < nm;
        off[m][0] = ok && 2 * p < count ? lbase + m * mstep : kOOB;
        off[m][1] = ok && 2 * p + 1 < count ? lbase + lstep + m * mstep : kOOB;
    }
    if constexpr (LAYOUT == 1) {
        // Columns: lines 2p and 2p+1 are adjacent floats (line stride 1), one
        // 8-byte load fetches both.  An odd last line of a tile reads the float
        // after its partner (inside the block: every domain block ends with a
        // spare float); gather_finish drops it.
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const float2 x = bload2(rmA, off[m][0]);
            R.ra[m][0] = x.x;
            R.ra[m][1] = x.y;
            const float2 y = hasB ? bload2(rmB, off[m][0]) : make_float2(0.f, 0.f);
            R.rb[m][0] = y.x;
            R.rb[m][1] = y.y;
        }
    } else if (hasB) {
#pragma unroll
        for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                R.ra[m][q] = bload(rmA, off[m][q]);
                R.rb[m][q] = bload(rmB, off[m][q]);
            }
    } else {
#pragma unroll
        for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                R.ra[m][q] = bload(rmA, off[m][q]);
                R.rb[m][q] = 0.f;
            }
    }
    // halos: item i of a side is (line ll, offset j), j in slots of 2^wlog >= w;
    // rows: j fastest, columns: ll fastest (coalesced either way)
    const Rsrc r1A = block_rsrc(A, td.s1_dom, td.s1_size);
    const Rsrc r1B = block_rsrc(B, td.s1_dom + (td.s1_b - td.s1_a), td.s1_size - (td.s1_b - td.s1_a));
    const Rsrc r2A = block_rsrc(A, td.s2_dom, td.s2_size);
    const Rsrc r2B = block_rsrc(B, td.s2_dom + (td.s2_b - td.s2_a), td.s2_size - (td.s2_b - td.s2_a));
    if constexpr (LAYOUT == 1) {
        // columns: item i is (pair pp, offset j), both lines of the pair (adjacent
        // floats) in one 8-byte load into registers 2k, 2k+1 (at most 2 items per
        // thread: P 2^wlog <= 2 NT); gather_finish drops a lone last line's value
        const int n_items = g.P << td.wlog;
        const int nk = (n_items + NT - 1) / NT;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            R.hpos1[k] = R.hpos2[k] = -1;
            R.h1a[k] = R.h1b[k] = R.h2a[k] = R.h2b[k] = 0.f;
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (k < nk) {
                const int i = threadIdx.x + k * NT;
                const int pp = i & (g.P - 1), j = i >> g.logP;
                const bool item = i < n_items && j < w;
                const bool ok = item && 2 * pp < count;
                const int line = td.line0 + 2 * pp;
                const int o1 = ok && td.s1_len >= 0
                                   ? (td.s1_a + (line + td.s1_line) * td.s1_ls + (td.s1_len - w - vel + j) * td.s1_ps) * 4
                                   : kOOB;
                const int o2 = ok && td.s2_len >= 0 ? (td.s2_a + (line + td.s2_line) * td.s2_ls + (j + vel) * td.s2_ps) * 4
                                                    : kOOB;
                const float2 x1 = bload2(r1A, o1), x2 = bload2(r2A, o2);
                R.h1a[2 * k] = x1.x;
                R.h1a[2 * k + 1] = x1.y;
                R.h2a[2 * k] = x2.x;
                R.h2a[2 * k + 1] = x2.y;
                if (hasB) {
                    const float2 y1 = bload2(r1B, o1), y2 = bload2(r2B, o2);
                    R.h1b[2 * k] = y1.x;
                    R.h1b[2 * k + 1] = y1.y;
                    R.h2b[2 * k] = y2.x;
                    R.h2b[2 * k + 1] = y2.y;
                }
                R.hpos1[k] = item ? g.stage_at(2 * pp, j) : -1;
                R.hpos2[k] = item ? g.stage_at(2 * pp, w + nm + j) : -1;
            }
        }
    } else {
        const int n_items = (2 * g.P) << td.wlog;
        const int nk = (n_items + NT - 1) / NT;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            R.hpos1[k] = R.hpos2[k] = -1;
            R.h1a[k] = R.h1b[k] = R.h2a[k] = R.h2b[k] = 0.f;
            if (k < nk) {
                const int i = threadIdx.x + k * NT;
                const int ll = i >> td.wlog;
                const int j = i & ((1 << td.wlog) - 1);
                const bool item = i < n_items && j < w;
                const bool ok = item && ll < count;
                const int line = td.line0 + ll;
                const int o1 = ok && td.s1_len >= 0
                                   ? (td.s1_a + (line + td.s1_line) * td.s1_ls + (td.s1_len - w - vel + j) * td.s1_ps) * 4
                                   : kOOB;
                const int o2 = ok && td.s2_len >= 0 ? (td.s2_a + (line + td.s2_line) * td.s2_ls + (j + vel) * td.s2_ps) * 4
                                                    : kOOB;
                R.h1a[k] = bload(r1A, o1);
                R.h2a[k] = bload(r2A, o2);
                if (hasB) {
                    R.h1b[k] = bload(r1B, o1);
                    R.h2b[k] = bload(r2B, o2);
                }
                R.hpos1[k] = item ? g.stage_at(ll, j) : -1;
                R.hpos2[k] = item ? g.stage_at(ll, w + nm + j) : -1;
            }
        }
    }
}

// Stages the raw values, then builds the windowed complex pairs.
template <int LAYOUT>
__device__ __forceinline__ void gather_finish(const StageArgs &a, const TileDesc &td, const TileGeo<LAYOUT> &g, int p,
                                              int t, const GatherRegs &R, float *stage, float *win, float2 (&v)[8]) {
    const int w = td.wlen, nm = td.n_main, vel = td.vel;
    __builtin_assume(w > 0 && w <= 32);
    const float c_s1 = td.c_s1, c_m1 = td.c_m1, c_s2 = td.c_s2, c_m2 = td.c_m2;
    // LDS addresses: line q of the pair at position pos = lb[q] + pos * ps
    const int ps = g.pos_step(), lb[2] = {g.line_base(2 * p), g.line_base(2 * p + 1)};
    const int tps = t * ps, mps = g.T * ps;
    const bool ok1 = LAYOUT == 0 || 2 * p + 1 < td.count;
    float raw[8][2];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int pos = t + m * g.T;
        const bool mid = pos >= w && pos < w + nm;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            raw[m][q] = q == 0 || ok1 ? R.ra[m][q] + R.rb[m][q] : 0.f;
            if (mid) stage[lb[q] + tps + m * mps] = raw[m][q];
        }
    }
    if constexpr (LAYOUT == 1) {
        // a pair's two halo values are adjacent in the staging (stage_at(2pp, j) even)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (R.hpos1[k] >= 0) {
                const bool two = (R.hpos1[k] & (2 * g.P - 1)) + 1 < td.count;
                *reinterpret_cast<float2 *>(&stage[R.hpos1[k]]) =
                    make_float2(R.h1a[2 * k] + R.h1b[2 * k], two ? R.h1a[2 * k + 1] + R.h1b[2 * k + 1] : 0.f);
                *reinterpret_cast<float2 *>(&stage[R.hpos2[k]]) =
                    make_float2(R.h2a[2 * k] + R.h2b[2 * k], two ? R.h2a[2 * k + 1] + R.h2b[2 * k + 1] : 0.f);
            }
        }
    } else {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (R.hpos1[k] >= 0) {
                stage[R.hpos1[k]] = R.h1a[k] + R.h1b[k];
                stage[R.hpos2[k]] = R.h2a[k] + R.h2b[k];
            }
        }
    }
    if ((int)threadIdx.x <= 2 * w) win[threadIdx.x] = R.win;
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int i = t + m * g.T;
        const bool left = i < w, right = i >= w + nm && i < 2 * w + nm, mid = i >= w && i < w + nm;
        const int j = i - w - nm;
        float out[2] = {0.f, 0.f};
        if (left || right) {
            const int mpos = left ? (vel ? w - i : w - 1 - i) : (vel ? nm - 2 - j : nm - 1 - j);
            const float cs = left ? c_s1 : c_s2, cm = left ? c_m1 : c_m2;
            const float wv = win[left ? i : w + 1 + j];
            const int ri = (w + mpos) * ps, si = tps + m * mps;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const float sv = stage[lb[q] + si];
                const float refl = stage[lb[q] + ri];
                out[q] = __builtin_fmaf(sv, cs, refl * cm) * wv;
            }
        }
        v[m] = make_float2(mid ? raw[m][0] : out[0], mid ? raw[m][1] : out[1]);
    }
}

// Forward transform, spectral derivative (the Hermitian filter h, which carries
// the 1/F of kernel/core/kernel_functions.cpp:243,306), inverse transform.
// LEAN: the filter and the twiddles are read from the tables (hsrc, tw) during
// the transforms instead of the registers preloaded by the caller (h, twr).
template <int F, int NT, int LAYOUT, bool LEAN>
__device__ __forceinline__ void derivative_fft(float2 (&v)[8], float2 (&h)[8], const float2 *twr, float2 *buf0,
                                               float2 *buf1, const float2 *hsrc, const float2 *tw) {
    constexpr int T = F / 8, P = NT / T;
    // (p, t) again from the thread index, with F known here
    const int tid = threadIdx.x;
    const int p = LAYOUT == 0 ? tid / T : tid % P, t = LAYOUT == 0 ? tid % T : tid / P;
    if constexpr (LEAN) {
        // filter and twiddles are loaded while the forward transform runs, the
        // twiddles one pass ahead (fft_run_ahead)
#pragma unroll
        for (int m = 0; m < 8; ++m) h[m] = hsrc[t + m * (F / 8)];
        const float2 w0[kTwPass] = {};
        fft_run_ahead<F, P, LAYOUT, false>(v, p, t, buf0, buf1, tw, w0);
    } else {
        fft_run<F, P, LAYOUT, false, false>(v, p, t, buf0, buf1, twr);
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = cmul(v[m], h[m]);
    // the inverse's first pass overwrites what the forward's last exchange read
    pass_sync<F, LAYOUT, true, 0>();
    if constexpr (LEAN) {
        const float2 w0[kTwPass] = {};
        fft_run_ahead<F, P, LAYOUT, true>(v, p, t, buf0, buf1, tw, w0);
    } else {
        fft_run<F, P, LAYOUT, true, false>(v, p, t, buf0, buf1, twr);
    }
}

constexpr int kMaxTw = 28;  // fft_tw_total(F) for F <= 8192

template <int F>
__device__ __forceinline__ void prefetch_twiddles(float2 *twr, int t, const float2 *tw) {
    static_assert(fft_tw_total(F) <= kMaxTw, "twiddle registers");
    fft_load_twiddles<F>(twr, t, tw);
}

__device__ __forceinline__ void track_max(const StageArgs &a, int scene, int cls, float m);

// Solver::update_field_values (kernel/Solver.cpp:174-181): out = t_n - s*L*coef,
// the last stage also t_n = out * PML factor.  Positions outside [0, n_out)
// and lines past the tile get an out-of-block offset (no store).
struct UpdateRegs {
    int off[8][2];
    float tnv[8][2], pm[8][2];
};

template <int LAYOUT>
__device__ __forceinline__ void prefetch_updates(const StageArgs &a, const TileDesc &td, const TileGeo<LAYOUT> &g,
                                                 int p, int t, UpdateRegs &u) {
    const int w = td.wlen, n_out = td.n_out;
    const bool ok0 = 2 * p < td.count, ok1 = 2 * p + 1 < td.count;
    const int pstep = td.o_ps * 4;
    const int base = (td.o_f + (td.line0 + 2 * p) * td.o_ls) * 4 + (t - w) * pstep, lstep = td.o_ls * 4,
              mstep = g.T * pstep;
    const bool use_pml = a.last && td.is_pml;
    const Rsrc rt = block_rsrc(a.tn, td.m_dom, td.m_size);
    const Rsrc rp = block_rsrc(a.pml, td.m_dom, td.m_size);
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int pos = t + m * g.T - w;
        const bool okp = pos >= 0 && pos < n_out;
        u.off[m][0] = okp && ok0 ? base + m * mstep : kOOB;
        u.off[m][1] = okp && ok1 ? base + lstep + m * mstep : kOOB;
        if constexpr (LAYOUT == 1) {
            // columns: both lines of the pair in one 8-byte access (as in gather_issue)
            const float2 x = bload2(rt, u.off[m][0]);
            const float2 f = use_pml ? bload2(rp, u.off[m][0]) : make_float2(1.f, 1.f);
            u.tnv[m][0] = x.x;
            u.tnv[m][1] = x.y;
            u.pm[m][0] = f.x;
            u.pm[m][1] = f.y;
        } else {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                u.tnv[m][q] = bload(rt, u.off[m][q]);
                u.pm[m][q] = use_pml ? bload(rp, u.off[m][q]) : 1.f;
            }
        }
    }
}

template <int LAYOUT>
__device__ __forceinline__ void finish_updates(const StageArgs &a, const TileDesc &td, int p, const float2 (&v)[8],
                                               const UpdateRegs &u) {
    const float sc = a.s * td.coef;
    const Rsrc ro = block_rsrc(a.out, td.m_dom, td.m_size);
    const Rsrc rt = block_rsrc(a.tn, td.m_dom, td.m_size);
    float nv[8][2];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        nv[m][0] = __builtin_fmaf(-sc, v[m].x, u.tnv[m][0]);
        nv[m][1] = __builtin_fmaf(-sc, v[m].y, u.tnv[m][1]);
    }
    if (a.umax) {
        float mx = 0.f;
#pragma unroll
        for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int q = 0; q < 2; ++q)
                if (u.off[m][q] != kOOB) mx = fmaxf(mx, fabsf(nv[m][q]));
        track_max(a, td.scene, td.vel ? 0 : 1, mx);
    }
    if (LAYOUT == 1 && 2 * p + 1 < td.count) {
        // columns, both lines in the tile: 8-byte stores
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            bstore2(ro, u.off[m][0], nv[m][0], nv[m][1]);
            if (a.last) bstore2(rt, u.off[m][0], nv[m][0] * u.pm[m][0], nv[m][1] * u.pm[m][1]);
        }
    } else {
#pragma unroll
        for (int m = 0; m < 8; ++m) {
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                bstore(ro, u.off[m][q], nv[m][q]);
                if (a.last) bstore(rt, u.off[m][q], nv[m][q] * u.pm[m][q]);
            }
        }
    }
}

// One derivative tile of one layout: NT threads, F = td.F in [FMIN, FMAX].
// Flavours differ only in when the loads that depend on nothing but the
// descriptor (filter, twiddles, t_n and PML operands) are issued:
//  kEarly: all before the gather completes (most registers);
//  kLate:  where they are used (fewest live registers, most resident waves);
//  kMid:   filter and twiddles up front (the transforms then never wait on
//          memory: the vector-memory counter is in order, so a late table
//          load waits for every load before it), t_n where used.
// All compute the same values in the same order.
constexpr int kEarly = 0, kLate = 1, kMid = 2;
template <int NT, int FMIN, int FMAX, int FL, int LAYOUT>
__device__ __forceinline__ void tile_layout(const StageArgs &a, const TileDesc &td, float2 *buf0, float2 *buf1) {
    const TileGeo<LAYOUT> g(td.F, NT);
    const int tid = threadIdx.x;
    const int p = g.pair(tid), t = g.lane_t(tid);
    GatherRegs R;
    gather_issue<NT, LAYOUT>(a, td, g, p, t, R);
    float2 h[8];
    constexpr bool LEAN = FL == kLate;
    float2 twr[LEAN ? 1 : kMaxTw];
    const float2 *tw = a.ctab + td.tw_off;
    const float2 *hsrc = a.ctab + td.h_off;
    if constexpr (!LEAN) {
#pragma unroll
        for (int m = 0; m < 8; ++m) h[m] = hsrc[t + m * g.T];
        switch (td.F) {
            case 128: if (FMIN <= 128 && FMAX >= 128) prefetch_twiddles<128>(twr, t, tw); break;
            case 256: if (FMIN <= 256 && FMAX >= 256) prefetch_twiddles<256>(twr, t, tw); break;
            case 512: if (FMIN <= 512 && FMAX >= 512) prefetch_twiddles<512>(twr, t, tw); break;
            case 1024: if (FMIN <= 1024 && FMAX >= 1024) prefetch_twiddles<1024>(twr, t, tw); break;
            case 2048: if (FMIN <= 2048 && FMAX >= 2048) prefetch_twiddles<2048>(twr, t, tw); break;
            case 4096: if (FMIN <= 4096 && FMAX >= 4096) prefetch_twiddles<4096>(twr, t, tw); break;
            case 8192: if (FMIN <= 8192 && FMAX >= 8192) prefetch_twiddles<8192>(twr, t, tw); break;
            default: break;
        }
    }
    UpdateRegs u;
    if constexpr (FL == kEarly) prefetch_updates<LAYOUT>(a, td, g, p, t, u);
    // buf1 holds the raw lines (2P lines, at most 18 NT floats); the window
    // has slots of its own (no pass of any pair writes them)
    __shared__ float win[2 * 32 + 1];
    float *stage = reinterpret_cast<float *>(buf1);
    float2 v[8];
    gather_finish<LAYOUT>(a, td, g, p, t, R, stage, win, v);
#define PSTD_FFT_CASE(FV)                                                                                \
    case FV:                                                                                             \
        if constexpr (FMIN <= FV && FMAX >= FV) derivative_fft<FV, NT, LAYOUT, LEAN>(v, h, twr, buf0, buf1, hsrc, tw); \
        break;
    switch (td.F) {
        PSTD_FFT_CASE(128)
        PSTD_FFT_CASE(256)
        PSTD_FFT_CASE(512)
        PSTD_FFT_CASE(1024)
        PSTD_FFT_CASE(2048)
        PSTD_FFT_CASE(4096)
        PSTD_FFT_CASE(8192)
        default: break;
    }
#undef PSTD_FFT_CASE
    if constexpr (FL != kEarly) prefetch_updates<LAYOUT>(a, td, g, p, t, u);
    finish_updates<LAYOUT>(a, td, p, v, u);
}

template <int NT, int FMIN, int FMAX, int FL>
__device__ __forceinline__ void derivative_tile(const StageArgs &a, const TileDesc &td, float2 *buf0, float2 *buf1) {
    if (td.dir == 0) tile_layout<NT, FMIN, FMAX, FL, 0>(a, td, buf0, buf1);
    else tile_layout<NT, FMIN, FMAX, FL, 1>(a, td, buf0, buf1);
}

// F = 8192 lines as two 4096-point transforms.  One line pair per 512-thread
// block; thread t holds positions n = t + 512 m (m = 0..15) of both lines (as
// one complex value z, line 0 real, line 1 imaginary).  One radix-2
// decimation-in-frequency step in registers: with W = e^{-2 pi i / 8192},
// a[n] = z[n] + z[n+4096] and b[n] = (z[n] - z[n+4096]) W^n are the inputs of
// the even and odd bins, Z[2k] = FFT4096(a)[k], Z[2k+1] = FFT4096(b)[k].  After
// the filter H (8192 bins, 1/8192 folded in) the inverse recombines in the same
// threads: y[n] = e[n] + W^-n o[n], y[n+4096] = e[n] - W^-n o[n], where
// e = IFFT4096(H[2k] Z[2k]) and o = IFFT4096(H[2k+1] Z[2k+1]).  Half the LDS
// of one 8192-point transform per block: two blocks share a CU.
constexpr int kSplitT = 512;
constexpr size_t kSplitLds = 2 * sizeof(float2) * LdsMap<4096, 1, 0>::SIZE;  // 72 KiB (the raw staging fits)

struct SplitGather {
    float ra[16][2], rb[16][2];
    float h1a[2], h1b[2], h2a[2], h2b[2];
    int hpos1, hpos2;
    float win;
};

// raw staging of the pair in LDS (floats): rows line-major at pitch 9F/8,
// columns position-major with the two lines adjacent
template <int LAYOUT>
__device__ __forceinline__ int split_stage_at(int ll, int pos) {
    return LAYOUT == 0 ? ll * (9 * 8192 / 8) + pos : ll + 2 * pos;
}

template <int LAYOUT>
__device__ __forceinline__ void split_gather_issue(const StageArgs &a, const TileDesc &td, int t, SplitGather &R) {
    const int w = td.wlen, nm = td.n_main, vel = td.vel, count = td.count;
    __builtin_assume(w > 0 && w <= 32);
    R.win = (int)threadIdx.x <= 2 * w ? a.ftab[td.win_off + threadIdx.x] : 0.f;
    const float *A = vel ? a.vsrc : a.p0a;
    const bool hasB = !vel && a.p0b != nullptr;
    const float *B = hasB ? a.p0b : A;
    const Rsrc rmA = block_rsrc(A, td.m_dom, td.m_size);
    const Rsrc rmB = block_rsrc(B, td.m_dom + (td.m_b - td.m_a), td.m_size - (td.m_b - td.m_a));
    const int pstep = td.m_ps * 4;
    const int lbase = (td.m_a + td.line0 * td.m_ls) * 4 + (t - w) * pstep, lstep = td.m_ls * 4,
              mstep = kSplitT * pstep;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int e = t + m * kSplitT - w;
        const bool ok = e >= 0 && e < nm;
        const int o0 = ok ? lbase + m * mstep : kOOB;
        const int o1 = ok && 1 < count ? lbase + lstep + m * mstep : kOOB;
        if constexpr (LAYOUT == 1) {
            // columns: the two lines are adjacent floats (an odd last line reads
            // the block's spare float after its partner; dropped below)
            const float2 x = bload2(rmA, o0);
            const float2 y = hasB ? bload2(rmB, o0) : make_float2(0.f, 0.f);
            R.ra[m][0] = x.x;
            R.ra[m][1] = x.y;
            R.rb[m][0] = y.x;
            R.rb[m][1] = y.y;
        } else {
            R.ra[m][0] = bload(rmA, o0);
            R.ra[m][1] = bload(rmA, o1);
            R.rb[m][0] = hasB ? bload(rmB, o0) : 0.f;
            R.rb[m][1] = hasB ? bload(rmB, o1) : 0.f;
        }
    }
    // halos: 2 w values per line from the side blocks; rows: item (line, j),
    // columns: item j with both lines in one 8-byte load
    const Rsrc r1A = block_rsrc(A, td.s1_dom, td.s1_size);
    const Rsrc r1B = block_rsrc(B, td.s1_dom + (td.s1_b - td.s1_a), td.s1_size - (td.s1_b - td.s1_a));
    const Rsrc r2A = block_rsrc(A, td.s2_dom, td.s2_size);
    const Rsrc r2B = block_rsrc(B, td.s2_dom + (td.s2_b - td.s2_a), td.s2_size - (td.s2_b - td.s2_a));
    const int i = threadIdx.x;
    const int ll = LAYOUT == 0 ? i >> td.wlog : 0, j = LAYOUT == 0 ? i & ((1 << td.wlog) - 1) : i;
    const bool item = LAYOUT == 0 ? (ll < 2 && j < w) : j < w;
    const bool ok = item && ll < count;
    const int line = td.line0 + ll;
    const int o1 = ok && td.s1_len >= 0
                       ? (td.s1_a + (line + td.s1_line) * td.s1_ls + (td.s1_len - w - vel + j) * td.s1_ps) * 4
                       : kOOB;
    const int o2 = ok && td.s2_len >= 0 ? (td.s2_a + (line + td.s2_line) * td.s2_ls + (j + vel) * td.s2_ps) * 4 : kOOB;
    if constexpr (LAYOUT == 1) {
        const float2 x1 = bload2(r1A, o1), x2 = bload2(r2A, o2);
        const float2 y1 = hasB ? bload2(r1B, o1) : make_float2(0.f, 0.f);
        const float2 y2 = hasB ? bload2(r2B, o2) : make_float2(0.f, 0.f);
        R.h1a[0] = x1.x; R.h1a[1] = x1.y; R.h2a[0] = x2.x; R.h2a[1] = x2.y;
        R.h1b[0] = y1.x; R.h1b[1] = y1.y; R.h2b[0] = y2.x; R.h2b[1] = y2.y;
    } else {
        R.h1a[0] = bload(r1A, o1);
        R.h2a[0] = bload(r2A, o2);
        R.h1b[0] = hasB ? bload(r1B, o1) : 0.f;
        R.h2b[0] = hasB ? bload(r2B, o2) : 0.f;
        R.h1a[1] = R.h2a[1] = R.h1b[1] = R.h2b[1] = 0.f;
    }
    R.hpos1 = item ? split_stage_at<LAYOUT>(ll, j) : -1;
    R.hpos2 = item ? split_stage_at<LAYOUT>(ll, w + nm + j) : -1;
}

// The windowed pair values at positions t + 512 m, combined into a and b.
template <int LAYOUT>
__device__ __forceinline__ void split_gather_finish(const TileDesc &td, int t, const SplitGather &R, float *stage,
                                                    float *win, const float2 *tw8, float2 (&va)[8], float2 (&vb)[8]) {
    const int w = td.wlen, nm = td.n_main, vel = td.vel;
    __builtin_assume(w > 0 && w <= 32);
    const float c_s1 = td.c_s1, c_m1 = td.c_m1, c_s2 = td.c_s2, c_m2 = td.c_m2;
    const bool ok1 = LAYOUT == 0 || 1 < td.count;
    float raw[16][2];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int pos = t + m * kSplitT;
        const bool mid = pos >= w && pos < w + nm;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            raw[m][q] = q == 0 || ok1 ? R.ra[m][q] + R.rb[m][q] : 0.f;
            if (mid) stage[split_stage_at<LAYOUT>(q, pos)] = raw[m][q];
        }
    }
    if (R.hpos1 >= 0) {
        if constexpr (LAYOUT == 1) {
            *reinterpret_cast<float2 *>(&stage[R.hpos1]) =
                make_float2(R.h1a[0] + R.h1b[0], ok1 ? R.h1a[1] + R.h1b[1] : 0.f);
            *reinterpret_cast<float2 *>(&stage[R.hpos2]) =
                make_float2(R.h2a[0] + R.h2b[0], ok1 ? R.h2a[1] + R.h2b[1] : 0.f);
        } else {
            stage[R.hpos1] = R.h1a[0] + R.h1b[0];
            stage[R.hpos2] = R.h2a[0] + R.h2b[0];
        }
    }
    if ((int)threadIdx.x <= 2 * w) win[threadIdx.x] = R.win;
    __syncthreads();
    // positions n and n + 4096 in turn, combined at once (fewer live registers)
    auto windowed = [&](int m) {
        const int i = t + m * kSplitT;
        const bool left = i < w, right = i >= w + nm && i < 2 * w + nm, mid = i >= w && i < w + nm;
        const int j = i - w - nm;
        float out[2] = {0.f, 0.f};
        if (left || right) {
            const int mpos = left ? (vel ? w - i : w - 1 - i) : (vel ? nm - 2 - j : nm - 1 - j);
            const float cs = left ? c_s1 : c_s2, cm = left ? c_m1 : c_m2;
            const float wv = win[left ? i : w + 1 + j];
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const float sv = stage[split_stage_at<LAYOUT>(q, i)];
                const float refl = stage[split_stage_at<LAYOUT>(q, w + mpos)];
                out[q] = __builtin_fmaf(sv, cs, refl * cm) * wv;
            }
        }
        return make_float2(mid ? raw[m][0] : out[0], mid ? raw[m][1] : out[1]);
    };
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const float2 z0 = windowed(m), z1 = windowed(m + 8);
        va[m] = cadd(z0, z1);
        vb[m] = cmul(csub(z0, z1), tw8[t + m * kSplitT]);
    }
}

// Update of the outputs at positions base + t + 512 m (out = t_n - s coef L;
// the last stage also t_n = out * PML), as finish_updates.
template <int LAYOUT>
__device__ __forceinline__ float split_update(const StageArgs &a, const TileDesc &td, int t, int base,
                                              const float2 (&y)[8]) {
    const int w = td.wlen, n_out = td.n_out;
    const bool ok1 = 1 < td.count;
    const int pstep = td.o_ps * 4;
    const int obase = (td.o_f + td.line0 * td.o_ls) * 4, lstep = td.o_ls * 4;
    const bool use_pml = a.last && td.is_pml;
    const Rsrc rt = block_rsrc(a.tn, td.m_dom, td.m_size);
    const Rsrc rp = block_rsrc(a.pml, td.m_dom, td.m_size);
    const Rsrc ro = block_rsrc(a.out, td.m_dom, td.m_size);
    const float sc = a.s * td.coef;
    int off[8];
    float tnv[8][2], pm[8][2];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int pos = base + t + m * kSplitT - w;
        off[m] = pos >= 0 && pos < n_out ? obase + pos * pstep : kOOB;
        if constexpr (LAYOUT == 1) {
            const float2 x = bload2(rt, off[m]);
            const float2 f = use_pml ? bload2(rp, off[m]) : make_float2(1.f, 1.f);
            tnv[m][0] = x.x; tnv[m][1] = x.y; pm[m][0] = f.x; pm[m][1] = f.y;
        } else {
            const int o1 = off[m] == kOOB || !ok1 ? kOOB : off[m] + lstep;
            tnv[m][0] = bload(rt, off[m]);
            tnv[m][1] = bload(rt, o1);
            pm[m][0] = use_pml ? bload(rp, off[m]) : 1.f;
            pm[m][1] = use_pml ? bload(rp, o1) : 1.f;
        }
    }
    float mx = 0.f;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const float n0 = __builtin_fmaf(-sc, y[m].x, tnv[m][0]);
        const float n1 = __builtin_fmaf(-sc, y[m].y, tnv[m][1]);
        if (off[m] != kOOB) mx = fmaxf(mx, ok1 ? fmaxf(fabsf(n0), fabsf(n1)) : fabsf(n0));
        if (LAYOUT == 1 && ok1) {
            bstore2(ro, off[m], n0, n1);
            if (a.last) bstore2(rt, off[m], n0 * pm[m][0], n1 * pm[m][1]);
        } else {
            const int o1 = off[m] == kOOB || !ok1 ? kOOB : off[m] + lstep;
            bstore(ro, off[m], n0);
            bstore(ro, o1, n1);
            if (a.last) {
                bstore(rt, off[m], n0 * pm[m][0]);
                bstore(rt, o1, n1 * pm[m][1]);
            }
        }
    }
    return mx;
}

template <int LAYOUT>
__device__ __forceinline__ void split_tile(const StageArgs &a, const TileDesc &td, float2 *lds) {
    const int t = threadIdx.x;
    float2 *buf0 = lds, *buf1 = lds + LdsMap<4096, 1, LAYOUT>::SIZE;
    __shared__ float win[2 * 32 + 1];
    SplitGather R;
    split_gather_issue<LAYOUT>(a, td, t, R);
    const float2 *tw8 = a.ctab + td.tw_off, *tw4 = a.ctab + td.tw2_off, *h = a.ctab + td.h_off;
    float2 va[8], vb[8];
    split_gather_finish<LAYOUT>(td, t, R, reinterpret_cast<float *>(lds), win, tw8, va, vb);
    const float2 w0[kTwPass] = {};
    // even bins: e = IFFT(H[2k] FFT(a)); the staging is read before the first pass writes
    __syncthreads();
    fft_run_ahead<4096, 1, LAYOUT, false>(va, 0, t, buf0, buf1, tw4, w0);
#pragma unroll
    for (int m = 0; m < 8; ++m) va[m] = cmul(va[m], h[2 * (t + m * kSplitT)]);
    __syncthreads();
    fft_run_ahead<4096, 1, LAYOUT, true>(va, 0, t, buf0, buf1, tw4, w0);
    // odd bins: o = IFFT(H[2k+1] FFT(b))
    __syncthreads();
    fft_run_ahead<4096, 1, LAYOUT, false>(vb, 0, t, buf0, buf1, tw4, w0);
#pragma unroll
    for (int m = 0; m < 8; ++m) vb[m] = cmul(vb[m], h[2 * (t + m * kSplitT) + 1]);
    __syncthreads();
    fft_run_ahead<4096, 1, LAYOUT, true>(vb, 0, t, buf0, buf1, tw4, w0);
    // y[n] = e[n] + W^-n o[n], y[n + 4096] = e[n] - W^-n o[n]
    float2 lo[8], hi[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        float2 wn = tw8[t + m * kSplitT];
        wn.y = -wn.y;
        const float2 wo = cmul(vb[m], wn);
        lo[m] = cadd(va[m], wo);
        hi[m] = csub(va[m], wo);
    }
    const float mx = fmaxf(split_update<LAYOUT>(a, td, t, 0, lo), split_update<LAYOUT>(a, td, t, 4096, hi));
    track_max(a, td.scene, td.vel ? 0 : 1, mx);
}

// One extra workgroup of a launch samples the receivers (p0 = px0 + py0 of the
// previous step, read from the arena `p0a`/`p0b` point at in the first stage).
__device__ void sample_receivers(const StageArgs &a) {
    if (a.slot < 0) return;
    for (int r = threadIdx.x; r < a.nrec; r += blockDim.x) {
        const ReceiverCell c = a.cells[r];
        if (c.px < 0) continue;  // spectral receiver
        a.samples[(size_t)a.slot * a.nrec + r] = a.p0a[c.px] + a.p0b[c.py];
    }
}

// Fields no segment updates keep their t_n value (openPSTD leaves their L at 0,
// kernel/Solver.cpp:174-181); the last stage applies the PML factor.  Each
// thread issues all of its loads before its stores.
template <int NT>
__device__ void passthrough_tile(const StageArgs &a, const TileDesc &td) {
    constexpr int PER = kPassChunk / NT;
    const Rsrc rt = block_rsrc(a.tn, td.m_dom, td.m_size);
    const Rsrc rp = block_rsrc(a.pml, td.m_dom, td.m_size);
    const Rsrc ro = block_rsrc(a.out, td.m_dom, td.m_size);
    const bool use_pml = a.last && td.is_pml;
    float v[PER], f[PER];
    int off[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int e = threadIdx.x + k * NT;
        off[k] = e < td.count ? (td.o_f + td.line0 + e) * 4 : kOOB;
        v[k] = bload(rt, off[k]);
        f[k] = use_pml ? bload(rp, off[k]) : 1.f;
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        bstore(ro, off[k], v[k]);
        if (use_pml) bstore(rt, off[k], v[k] * f[k]);
    }
    if (a.umax) {
        float mx = 0.f;
#pragma unroll
        for (int k = 0; k < PER; ++k) mx = fmaxf(mx, fabsf(v[k]));
        track_max(a, td.scene, td.vel ? 0 : 1, mx);  // vel: 1 for px/py (as the velocity lines that write them)
    }
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b runs on XCD
// b % 8) and every XCD has its own L2.  With `remap`, XCD x works through a
// contiguous run of the tile list instead of every 8th tile, so tiles of
// neighbouring lines (the columns of one cache line, the rows of one halo)
// share that L2.  A bijection of [0, n).
__device__ __forceinline__ int tile_index(int b, int n, int remap) {
    if (!remap) return b;
    constexpr int X = 8;
    const int x = b % X, l = b / X, q = n / X, r = n % X;
    return x * q + min(x, r) + l;
}

// F <= 2048 class, NT = 128, 256 or 512 threads (128: F <= 1024 only): two LDS
// buffers of P * F = 8 * NT complex values each (padded).  Wider blocks hold
// more lines per tile (for columns, more of each row per wave access) at the
// same waves per CU; narrower ones spread a small scene over more CUs.
template <int NT, int FL>
__global__ void __launch_bounds__(NT) stage_small_kernel(StageArgs a) {
    constexpr int kBuf = lds_padded(8 * NT);
    __shared__ float2 lds[2 * kBuf];
    // a copy taken before any store: the descriptor stays in scalar registers
    const TileDesc td = a.tiles[tile_index(blockIdx.x, gridDim.x, a.remap)];
    if (td.kind == 2) {
        sample_receivers(a);
        return;
    }
    if (td.kind == 1) {
        passthrough_tile<NT>(a, td);
        return;
    }
    derivative_tile<NT, 128, (NT >= 256 ? 2048 : 1024), FL>(a, td, lds, lds + kBuf);
}

template <int F, int FL>
__global__ void __launch_bounds__(F / 8) stage_big_kernel(StageArgs a) {
    extern __shared__ float2 lds_dyn[];
    const TileDesc td = a.tiles[tile_index(blockIdx.x, gridDim.x, a.remap)];
    derivative_tile<F / 8, F, F, FL>(a, td, lds_dyn, lds_dyn + LdsMap<F, 1, 0>::SIZE);
}

// two blocks per CU: 16 waves, at most 128 VGPRs each
__global__ void __launch_bounds__(kSplitT) __attribute__((amdgpu_waves_per_eu(4))) stage_split_kernel(StageArgs a) {
    extern __shared__ float2 lds_split[];
    const TileDesc td = a.tiles[tile_index(blockIdx.x, gridDim.x, a.remap)];
    if (td.dir == 0) split_tile<0>(a, td, lds_split);
    else split_tile<1>(a, td, lds_split);
}

__global__ void receiver_kernel(const float *__restrict__ arena, const ReceiverCell *__restrict__ cells, int nrec,
                                float *__restrict__ samples, int slot) {
    // p0 = px0 + py0 (kernel/Solver.cpp:114-116)
    for (int r = threadIdx.x; r < nrec; r += blockDim.x) {
        const ReceiverCell c = cells[r];
        if (c.px < 0) continue;  // spectral receiver: sampled by the interpolation passes
        samples[(size_t)slot * nrec + r] = arena[c.px] + arena[c.py];
    }
}

// Spectral-interpolation receivers (Receiver::compute_with_si,
// kernel/core/Receiver.cpp:85-121), two passes after each step.  The
// fractional shift of one pressure line, read at one output, is a fixed linear
// function of the line's K = n + 2w raw values [side-1 tail | main | side-2
// head] (p0 = px + py): the row of the line operator (shift filter composed
// with the window and reflections) built in double on the host.  Pass 1: one
// wave per line of the domains the receiver's Y line reads, the dot product
// of its values with its X-shift row, into the receiver's column buffer
// [side-1 tail | domain rows | side-2 head].  Pass 2: one wave per receiver,
// the column buffer against the Y-shift row of the receiver's row.
struct SILine {
    long long m_px, m_py;    // arena offsets of the line's first main value in px0, py0
    long long s1_px, s1_py;  // first of the w side-1 tail values; -1: no neighbour (zeros)
    long long s2_px, s2_py;  // first of the w side-2 head values; -1: zeros
    int n_main, w;
    int coef;                // first of the line's K weights in the weight table
    int dst;                 // element of the column buffer
};

struct SIPoint {
    int coef, K;             // Y-shift weights of the receiver's row
    int col;                 // first element of the receiver's column buffer
    int rec;                 // receiver index (sample position)
};

__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ __launch_bounds__(256) void si_line_kernel(const float *__restrict__ arena, const SILine *__restrict__ lines,
                                                      int nlines, const float *__restrict__ coef,
                                                      float *__restrict__ col) {
    // jobs are sorted by line: the receivers reading one line are neighbours,
    // and XCD-contiguous runs of blocks put them on one XCD (its L2 serves the
    // line after the first read)
    const int lane = threadIdx.x & 63, j = tile_index(blockIdx.x, gridDim.x, 1) * 4 + (threadIdx.x >> 6);
    if (j >= nlines) return;
    const SILine L = lines[j];
    const int w = L.w, n = L.n_main, K = n + 2 * w;
    const float *cf = coef + L.coef;
    float acc = 0.f;
    for (int k = lane; k < K; k += 64) {
        float u = 0.f;
        if (k < w) {
            if (L.s1_px >= 0) u = arena[L.s1_px + k] + arena[L.s1_py + k];
        } else if (k < w + n) {
            u = arena[L.m_px + (k - w)] + arena[L.m_py + (k - w)];
        } else if (L.s2_px >= 0) {
            u = arena[L.s2_px + (k - w - n)] + arena[L.s2_py + (k - w - n)];
        }
        acc = fmaf(cf[k], u, acc);
    }
    acc = wave_sum(acc);
    if (lane == 0) col[L.dst] = acc;
}

__global__ __launch_bounds__(64) void si_point_kernel(const float *__restrict__ col, const SIPoint *__restrict__ pts,
                                                      const float *__restrict__ coef, float *__restrict__ samples,
                                                      int nrec, int slot) {
    const SIPoint P = pts[blockIdx.x];
    float acc = 0.f;
    for (int k = threadIdx.x; k < P.K; k += 64) acc = fmaf(coef[P.coef + k], col[P.col + k], acc);
    acc = wave_sum(acc);
    if (threadIdx.x == 0) samples[(size_t)slot * nrec + P.rec] = acc;
}

struct P0Range {
    long long px, py;  // arena element offsets of px0 and py0 of one domain
    long long n;
};

__global__ void p0_materialise_kernel(const float *__restrict__ src, float *__restrict__ p0i,
                                      const P0Range *__restrict__ ranges, int nr) {
    // P0I.px-slot = px0 + py0 of every domain
    for (int k = blockIdx.y; k < nr; k += gridDim.y) {
        const P0Range r = ranges[k];
        for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < r.n; e += (long long)gridDim.x * blockDim.x)
            p0i[r.px + e] = src[r.px + e] + src[r.py + e];
    }
}

// ------------------------------------------------------------- GEMM form
// The windowed line derivative of a segment is a fixed linear operator: the
// dense matrix D (n_out x K) of plan_tables.h, applied to the raw line
// u = [side1 tail | main | side2 head] (K values).  One workgroup (4 waves,
// two per CU) computes 128 lines x 128 outputs of C = U D^T on MFMA and
// applies the RK update:
//  * the input lines are read as FP32 straight from the stage state (p0 as
//    px + py), scaled by the scene's power of two 2^e and split into two
//    16-bit parts while they are staged into LDS (U = hi + lo: FP16 keeps 22
//    bits when every scaled value is below 2^15, BF16 16 bits at any scale);
//  * D is stored split the same way (one power of two per matrix), as the
//    LDS images of its 32-k chunks, copied straight from memory into LDS;
//  * C = U_hi D_hi + U_hi D_lo + U_lo D_hi with v_mfma_f32_32x32x16_{f16,bf16},
//    FP32 accumulation, over k-chunks of 32 through double-buffered LDS (one
//    workgroup barrier per chunk; the loads of chunk c+1 are in flight while
//    chunk c is multiplied);
//  * the scene's scale comes from the largest |value| every kernel of the
//    previous stage wrote (a per-scene atomic maximum in three rotating
//    slots: stage k reads slot (k+2)%3, writes slot k%3 and clears slot
//    (k+1)%3), so no pass over the inputs precedes the multiplication;
//  * pressure lines have n+1 outputs: the multiplication covers the first n,
//    the last is an FP32 dot product of the raw line with D's last row,
//    accumulated by the threads while they stage the line (its 128-wide
//    output tiles then fit n exactly instead of wasting most of a tile).
// X lines put U on the MFMA A side (rows of C = lines, columns = outputs),
// Y lines put D there (rows = outputs, columns = lines), so that consecutive
// lanes store consecutive floats in both directions.
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <typename E>
struct MfmaOps;
template <>
struct MfmaOps<_Float16> {
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    typedef _Float16 v4 __attribute__((ext_vector_type(4)));
    __device__ static f32x16 mfma(v8 a, v8 b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }
};
template <>
struct MfmaOps<__bf16> {
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    typedef __bf16 v4 __attribute__((ext_vector_type(4)));
    __device__ static f32x16 mfma(v8 a, v8 b, f32x16 c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
};

constexpr int kGemmThreads = 256;  // 4 waves: 2 line groups x 2 output groups of 64 x 64
constexpr int kGL = 128;           // lines per tile
constexpr int kGN = 128;           // outputs per tile
// (line, 8-k group) pieces of a chunk each thread stages: 2
constexpr int kGSlots = kGL * 32 / 8 / kGemmThreads;
constexpr int kGK = 32;            // k per LDS chunk (two MFMA k-steps)
constexpr int kGP = kGK + 8;       // LDS row pitch in 2-byte elements (80 B: conflict-free 16-byte operand reads)
constexpr int kGemmLdsU = 2 * 2 * kGL * kGP;  // elements: [buffer][part][line][k]
constexpr int kGemmLdsD = 2 * 2 * kGN * kGP;  // elements: [buffer][part][output][k]
constexpr size_t kGemmLdsBytes = (size_t)(kGemmLdsU + kGemmLdsD) * 2;  // 80 KiB: two workgroups per CU
constexpr int kGemmDChunk = 2 * kGN * kGP;  // elements of one chunk's D image (both parts)
static_assert(kGemmDChunk * 2 == 5 * 1024 * (kGemmThreads / 64), "D image: 5 KiB-loads per wave");
static_assert(kGSlots == 2, "staging map assumes two slots per thread");

struct GemmTile {
    TileDesc t;            // segment addressing; line0/count: the tile's lines (count <= 128)
    // D of the tile's 128 outputs as LDS images (2-byte elements): chunk c's
    // image [part hi, lo][128 outputs][kGP] (k of the chunk, then zero pad) at
    // element d_img + c * kGemmDChunk; d_len elements in all
    long long d_img, d_len;
    int n_rows, k_pad;     // rows of D (multiple of kGN) and columns (multiple of kGK)
    int n0;                // first output of the tile
    int n_g;               // outputs the multiplication produces (n_out, or n_out - 1 with the dot row)
    int d_exp;             // D was stored as D * 2^d_exp
    int K;                 // line length
    int dot_off;           // >= 0: ftab offset of D's last row (FP32): this tile writes output n_out - 1
};

// Per-scene maxima of |value| written in a stage: umax[(slot * n_scenes + scene) * 2 + class],
// class 0 the pressure fields (px, py), class 1 the velocities; float bits.
__device__ __forceinline__ void track_max(const StageArgs &a, int scene, int cls, float m) {
    if (!a.umax) return;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0 && m > 0.f)
        atomicMax(&a.umax[((size_t)a.slot_w * a.n_scenes + scene) * 2 + cls], __float_as_uint(m));
}

// Power of two that brings a line whose values are bounded by m below 2^15
// (FP16 keeps both parts normal down to 2^-3 of that bound).
__device__ __forceinline__ int gemm_scale_exp(const StageArgs &a, int scene, int vel) {
    const unsigned bits = a.umax[((size_t)a.slot_r * a.n_scenes + scene) * 2 + (vel ? 1 : 0)];
    const float m = __uint_as_float(bits) * (vel ? 1.f : 2.f);  // p0 = px + py
    return m > 0.f ? min(max(14 - ilogbf(m), -100), 100) : 0;
}

// Raw input values of one chunk held by a thread: 8 consecutive k of one line
// (X lines: 4 threads per line, so 4 lanes read 128 contiguous bytes of a row;
// Y lines: one line per lane, so a wavefront reads 64 adjacent columns of a
// row).  The second array of p0 = px + py is kept apart and added when the
// chunk is converted, so that nothing waits for the loads of chunk c+2 before
// the MFMAs of chunk c; dk: the FP32 dot row's coefficients of these k.
struct UChunk {
    float a[kGSlots][8], b[kGSlots][8], dk[kGSlots][8];
};

// Staging map of slot q: X lines 4 threads per line (lines tid/4 and tid/4 +
// 64); Y lines one line per lane, the k-group that of the whole wavefront
// (groups tid/128 and tid/128 + 2).
template <int DIR>
__device__ __forceinline__ int stage_line(int tid, int q) {
    return DIR == 0 ? (tid >> 2) + q * (kGemmThreads / 4) : tid & (kGL - 1);
}
template <int DIR>
__device__ __forceinline__ int stage_kq(int tid, int q) {
    return DIR == 0 ? (tid & 3) * 8 : __builtin_amdgcn_readfirstlane((tid >> 7) + q * (kGemmThreads / kGL)) * 8;
}

// raw buffer over `bytes` bytes at byte offset `byte_off` of `base`
__device__ __forceinline__ Rsrc bytes_rsrc(const void *base, long long byte_off, int bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)((const char *)base + byte_off), (short)0, bytes, 0x00020000);
}

__device__ __forceinline__ float4 bload4(Rsrc r, int off_bytes) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off_bytes, 0, 0));
}

// Line `line` (absolute) of a tile: byte offsets of position k of its raw
// input line in the three blocks it comes from (main, side1, side2).
struct LineAddr {
    int w, nm, vel, K;
    int m_base, m_ps, s1_base, s1_ps, s2_base, s2_ps;  // offset of k = 0 of each piece, byte step per k
    bool s1, s2;
    __device__ LineAddr(const TileDesc &td, int line, int K_) {
        w = td.wlen;
        nm = td.n_main;
        vel = td.vel;
        K = K_;
        m_ps = td.m_ps * 4;
        m_base = (td.m_a + line * td.m_ls) * 4 - w * m_ps;
        s1 = td.s1_len >= 0;
        s2 = td.s2_len >= 0;
        s1_base = (td.s1_a + (line + td.s1_line) * td.s1_ls + (td.s1_len - w - vel) * td.s1_ps) * 4;
        s1_ps = td.s1_ps * 4;
        s2_base = (td.s2_a + (line + td.s2_line) * td.s2_ls + (vel - w - nm) * td.s2_ps) * 4;
        s2_ps = td.s2_ps * 4;
    }
    __device__ int main_off(int k) const { return k >= w && k < w + nm ? m_base + k * m_ps : kOOB; }
    __device__ int s1_off(int k) const { return s1 && k < w ? s1_base + k * s1_ps : kOOB; }
    __device__ int s2_off(int k) const { return s2 && k >= w + nm && k < K ? s2_base + k * s2_ps : kOOB; }
};

// The arrays a tile's lines are read from: A (vx / vy for velocity lines, px
// or the explicit p0 for pressure lines) and, for p0 = px + py, B (py).  The
// three pieces of a line differ only in the domain block.
struct GemmSrc {
    const float *A, *B;
    bool hasB;
};

__device__ __forceinline__ GemmSrc gemm_src(const StageArgs &a, const TileDesc &td) {
    GemmSrc r;
    r.A = td.vel ? a.vsrc : a.p0a;
    r.hasB = !td.vel && a.p0b != nullptr;
    r.B = r.hasB ? a.p0b : r.A;
    return r;
}

// Which piece of the line a whole chunk [k0, k0 + kGK) lies in (the same for
// every line of a tile).  kPieceS2 chunks may run past K: positions from K on
// read zero (a value from past the halo could be another field, outside the
// range the scene's scale was chosen for).
constexpr int kPieceMain = 0, kPieceS1 = 1, kPieceS2 = 2, kPieceMixed = 3, kPieceNone = 4;
__device__ __forceinline__ int chunk_piece(const TileDesc &td, int K, int k0) {
    const int w = td.wlen, nm = td.n_main;
    return k0 >= K                              ? kPieceNone
           : k0 >= w && k0 + kGK <= w + nm      ? kPieceMain
           : k0 + kGK <= w                      ? kPieceS1
           : k0 >= w + nm                       ? kPieceS2
                                                : kPieceMixed;
}

// Buffer resources (A, B arrays) of one piece's domain block.
// (Selections are between values, never between struct members: a select of
// member lvalues keeps the struct in scratch memory.)
template <class T>
__device__ __forceinline__ T pick3(int piece, T main, T s1, T s2) {
    return piece == kPieceMain ? main : piece == kPieceS1 ? s1 : s2;
}
__device__ __forceinline__ void piece_rsrc(const GemmSrc &src, const TileDesc &td, int piece, Rsrc &rA, Rsrc &rB) {
    const long long dom = pick3<long long>(piece, td.m_dom, td.s1_dom, td.s2_dom);
    const int size = pick3<int>(piece, td.m_size, td.s1_size, td.s2_size);
    const int ab = pick3<int>(piece, td.m_b - td.m_a, td.s1_b - td.s1_a, td.s2_b - td.s2_a);
    rA = block_rsrc(src.A, dom, size);
    rB = block_rsrc(src.B, dom + ab, size - ab);
}

template <int DIR, bool HASB>
__device__ __forceinline__ void gemm_load_u1(const GemmSrc &src, const TileDesc &td, const LineAddr &L, bool ok, int k0,
                                             int q, Rsrc dotr, bool dot, UChunk &u) {
    const int kb = k0 + stage_kq<DIR>(threadIdx.x, q);
    const int piece = chunk_piece(td, L.K, k0);
    if (piece == kPieceMixed) {
        // a chunk across a piece boundary: each value from whichever block holds it
        Rsrc mA, mB, s1A, s1B, s2A, s2B;
        piece_rsrc(src, td, kPieceMain, mA, mB);
        piece_rsrc(src, td, kPieceS1, s1A, s1B);
        piece_rsrc(src, td, kPieceS2, s2A, s2B);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = kb + j;
            const int om = ok ? L.main_off(k) : kOOB, o1 = ok ? L.s1_off(k) : kOOB, o2 = ok ? L.s2_off(k) : kOOB;
            u.a[q][j] = bload(mA, om) + bload(s1A, o1) + bload(s2A, o2);
            u.b[q][j] = HASB ? bload(mB, om) + bload(s1B, o1) + bload(s2B, o2) : 0.f;
        }
    } else {
        // one block for the whole chunk
        Rsrc rA, rB;
        piece_rsrc(src, td, piece, rA, rB);
        const bool in = ok && (piece == kPieceMain || (piece == kPieceS1 && L.s1) || (piece == kPieceS2 && L.s2));
        const int base = pick3<int>(piece, L.m_base, L.s1_base, L.s2_base);
        const int ps = pick3<int>(piece, L.m_ps, L.s1_ps, L.s2_ps);
        const int off = in && piece != kPieceNone ? base + kb * ps : kOOB;
        if (DIR == 0 && piece != kPieceS2) {
            // X lines inside a row: two 16-byte loads per array
            const float4 x0 = bload4(rA, off), x1 = bload4(rA, off == kOOB ? kOOB : off + 16);
            u.a[q][0] = x0.x; u.a[q][1] = x0.y; u.a[q][2] = x0.z; u.a[q][3] = x0.w;
            u.a[q][4] = x1.x; u.a[q][5] = x1.y; u.a[q][6] = x1.z; u.a[q][7] = x1.w;
            if constexpr (HASB) {
                const float4 y0 = bload4(rB, off), y1 = bload4(rB, off == kOOB ? kOOB : off + 16);
                u.b[q][0] = y0.x; u.b[q][1] = y0.y; u.b[q][2] = y0.z; u.b[q][3] = y0.w;
                u.b[q][4] = y1.x; u.b[q][5] = y1.y; u.b[q][6] = y1.z; u.b[q][7] = y1.w;
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) u.b[q][j] = 0.f;
            }
        } else {
            // Y lines (a wavefront reads a row segment per k), or the side-2
            // chunks of X lines (they may run to the end of the block): 4-byte loads
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int o = off == kOOB || kb + j >= L.K ? kOOB : off + j * ps;
                u.a[q][j] = bload(rA, o);
                u.b[q][j] = HASB ? bload(rB, o) : 0.f;
            }
        }
    }
    if (dot) {
#pragma unroll
        for (int j = 0; j < 8; ++j) u.dk[q][j] = bload(dotr, (kb + j) * 4);  // zero from K on
    }
}

template <int DIR, bool HASB>
__device__ __forceinline__ void gemm_load_u(const GemmSrc &src, const TileDesc &td, const LineAddr (&L)[kGSlots],
                                            const bool (&ok)[kGSlots], int k0, Rsrc dotr, bool dot, UChunk &u) {
#pragma unroll
    for (int q = 0; q < kGSlots; ++q) gemm_load_u1<DIR, HASB>(src, td, L[q], ok[q], k0, q, dotr, dot, u);
}

// Main-piece addressing of a tile's lines: every k of a main chunk lies in the
// line's own block, so the byte offset splits into a per-lane part (line and
// k-group; kOOB for lanes past the tile's lines) and a uniform part (the
// chunk's k, passed as the buffer instruction's scalar offset).  Both parts
// are non-negative.
struct MainU {
    Rsrc rA, rB;
    int voff[kGSlots];  // per lane and slot
    int ps;    // bytes per k
    int kofs;  // k of the first main position (the line's halo length)
};

template <int DIR, bool HASB>
__device__ __forceinline__ MainU main_u(const GemmSrc &src, const TileDesc &td, const bool (&ok)[kGSlots]) {
    MainU m;
    m.rA = block_rsrc(src.A, td.m_dom, td.m_size);
    m.rB = block_rsrc(src.B, td.m_dom + (td.m_b - td.m_a), td.m_size - (td.m_b - td.m_a));
    m.ps = td.m_ps * 4;
    m.kofs = td.wlen;
#pragma unroll
    for (int q = 0; q < kGSlots; ++q) {
        const int line = td.line0 + stage_line<DIR>(threadIdx.x, q);
        m.voff[q] = ok[q] ? (td.m_a + line * td.m_ls) * 4 + stage_kq<DIR>(threadIdx.x, q) * m.ps : kOOB;
    }
    return m;
}

template <int DIR, bool HASB>
__device__ __forceinline__ void gemm_load_u_main(const MainU &m, int k0, Rsrc dotr, bool dot, UChunk &u) {
    const int kk = k0 - m.kofs;  // >= 0 in a main chunk
#pragma unroll
    for (int q = 0; q < kGSlots; ++q) {
        if constexpr (DIR == 0) {
            // X lines (element stride 1): two 16-byte loads per array
            const int so = kk * 4;
            const float4 x0 = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(m.rA, m.voff[q], so, 0));
            const float4 x1 =
                __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(m.rA, m.voff[q] + 16, so, 0));
            u.a[q][0] = x0.x; u.a[q][1] = x0.y; u.a[q][2] = x0.z; u.a[q][3] = x0.w;
            u.a[q][4] = x1.x; u.a[q][5] = x1.y; u.a[q][6] = x1.z; u.a[q][7] = x1.w;
            if constexpr (HASB) {
                const float4 y0 =
                    __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(m.rB, m.voff[q], so, 0));
                const float4 y1 =
                    __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(m.rB, m.voff[q] + 16, so, 0));
                u.b[q][0] = y0.x; u.b[q][1] = y0.y; u.b[q][2] = y0.z; u.b[q][3] = y0.w;
                u.b[q][4] = y1.x; u.b[q][5] = y1.y; u.b[q][6] = y1.z; u.b[q][7] = y1.w;
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) u.b[q][j] = 0.f;
            }
        } else {
            // Y lines: a wavefront reads 64 adjacent columns of a row per k
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int so = (kk + j) * m.ps;
                u.a[q][j] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(m.rA, m.voff[q], so, 0));
                u.b[q][j] =
                    HASB ? __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(m.rB, m.voff[q], so, 0)) : 0.f;
            }
        }
        if (dot) {
            const int kb = k0 + stage_kq<DIR>(threadIdx.x, q);
#pragma unroll
            for (int j = 0; j < 8; ++j) u.dk[q][j] = bload(dotr, (kb + j) * 4);
        }
    }
}

// Copies chunk c of the tile's D image into LDS buffer `dbuf` (20 KiB, straight
// from memory into LDS: 20 wavefront loads of 1 KiB, 5 per wave).
__device__ __forceinline__ void gemm_dma_d(const Rsrc &dr, void *dbuf, int c) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int j = 0; j < kGemmDChunk * 2 / 1024 / (kGemmThreads / 64); ++j)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(dr, (char *)dbuf + (j * (kGemmThreads / 64) + w) * 1024, 16,
                                                 lane * 16 + w * 1024, c * kGemmDChunk * 2 + j * 4096, 0, 0);
}

// Converts and writes one chunk into LDS buffer `buf` (8 k of one line per
// thread as one 16-byte write per part); accumulates the FP32 dot product of
// the last output row (dot).
template <typename E, int DIR, bool HASB>
__device__ __forceinline__ void gemm_stage(E *ulds, int buf, const UChunk &u, float sc, bool dot,
                                           float (&gv)[kGSlots]) {
    using V8 = typename MfmaOps<E>::v8;
    const int tid = threadIdx.x;
    E *U0 = ulds + (buf * 2 + 0) * kGL * kGP, *U1 = ulds + (buf * 2 + 1) * kGL * kGP;
#pragma unroll
    for (int q = 0; q < kGSlots; ++q) {
        const int row = stage_line<DIR>(tid, q), kq = stage_kq<DIR>(tid, q);
        V8 hi, lo;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float raw = HASB ? u.a[q][j] + u.b[q][j] : u.a[q][j];
            const float x = raw * sc;
            const E h = (E)x;
            hi[j] = h;
            lo[j] = (E)(x - (float)h);
        }
        *reinterpret_cast<V8 *>(U0 + row * kGP + kq) = hi;
        *reinterpret_cast<V8 *>(U1 + row * kGP + kq) = lo;
    }
    // the dot row (a uniform branch: only the last output tile of a line block)
    if (dot) {
#pragma unroll
        for (int q = 0; q < kGSlots; ++q)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                gv[q] = __builtin_fmaf(HASB ? u.a[q][j] + u.b[q][j] : u.a[q][j], u.dk[q][j], gv[q]);
    }
}

// The MFMAs of one chunk (LDS buffer `buf`): wave (lg, og) accumulates lines
// 64 lg .. +63 x outputs 64 og .. +63 (2 x 2 blocks of 32 x 32).
template <typename E, int DIR>
__device__ __forceinline__ void gemm_chunk(const E *ulds, const E *dlds, int buf, int lg, int og, f32x16 (&acc)[2][2]) {
    using V8 = typename MfmaOps<E>::v8;
    const int lane = threadIdx.x & 63;
    const E *U0 = ulds + (buf * 2 + 0) * kGL * kGP, *U1 = ulds + (buf * 2 + 1) * kGL * kGP;
    const E *D0 = dlds + (buf * 2 + 0) * kGN * kGP, *D1 = dlds + (buf * 2 + 1) * kGN * kGP;
    // rows of C from R (U for X, D for Y) at r_base, columns from Q at q_base
    const E *R0 = DIR == 0 ? U0 : D0, *R1 = DIR == 0 ? U1 : D1;
    const E *Q0 = DIR == 0 ? D0 : U0, *Q1 = DIR == 0 ? D1 : U1;
    const int r_base = DIR == 0 ? 64 * lg : 64 * og, q_base = DIR == 0 ? 64 * og : 64 * lg;
#pragma unroll
    for (int s = 0; s < kGK / 16; ++s) {
        const int kk = s * 16 + (lane >> 5) * 8;
        V8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int rr = (r_base + 32 * i + (lane & 31)) * kGP + kk, qq = (q_base + 32 * i + (lane & 31)) * kGP + kk;
            ah[i] = *reinterpret_cast<const V8 *>(R0 + rr);
            al[i] = *reinterpret_cast<const V8 *>(R1 + rr);
            bh[i] = *reinterpret_cast<const V8 *>(Q0 + qq);
            bl[i] = *reinterpret_cast<const V8 *>(Q1 + qq);
        }
        // each product over the four independent accumulators in turn (no
        // back-to-back MFMAs on one accumulator)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = MfmaOps<E>::mfma(ah[i], bh[j], acc[i][j]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = MfmaOps<E>::mfma(ah[i], bl[j], acc[i][j]);
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = MfmaOps<E>::mfma(al[i], bh[j], acc[i][j]);
    }
}

template <typename E, int DIR, bool HASB>
__device__ __forceinline__ void gemm_tile(const StageArgs &a, const GemmTile &gt, const float *ftab,
                                          unsigned char *gemm_lds) {
    E *ulds = reinterpret_cast<E *>(gemm_lds);
    E *dlds = ulds + kGemmLdsU;
    const TileDesc &td = gt.t;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lg = wave >> 1, og = wave & 1;
    const int e = gemm_scale_exp(a, td.scene, td.vel);
    const float sc = ldexpf(1.f, e);
    const GemmSrc r = gemm_src(a, td);
    const Rsrc dr = bytes_rsrc(a.dmat, gt.d_img * 2, (int)(gt.d_len * 2));
    const bool dot = gt.dot_off >= 0;
    const Rsrc dotr = bytes_rsrc(ftab, dot ? (long long)gt.dot_off * 4 : 0, dot ? gt.K * 4 : 0);
    const int nch = gt.k_pad / kGK;
    bool line_ok[kGSlots];
#pragma unroll
    for (int q = 0; q < kGSlots; ++q) line_ok[q] = stage_line<DIR>(tid, q) < td.count;
    const LineAddr L[kGSlots] = {LineAddr(td, td.line0 + stage_line<DIR>(tid, 0), gt.K),
                                 LineAddr(td, td.line0 + stage_line<DIR>(tid, 1), gt.K)};
    // waves whose lines or outputs lie outside the tile skip their MFMAs
    const bool active = 64 * lg < td.count && 64 * og < gt.n_g - gt.n0;
    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[i][j][q] = 0.f;
    float gv[kGSlots] = {0.f, 0.f};
    // One register set: the loads of chunk c+1 are issued right after the
    // barrier that makes chunk c visible, and are in flight while the MFMAs of
    // chunk c run; then chunk c+1 is converted into the other LDS buffer.
    UChunk u;
    gemm_load_u<DIR, HASB>(r, td, L, line_ok, 0, dotr, dot, u);
    gemm_dma_d(dr, dlds, 0);
    gemm_stage<E, DIR, HASB>(ulds, 0, u, sc, dot, gv);
    __syncthreads();
    // Chunks whose every k lies in the main piece, [cm0, cm1), load through
    // MainU (no per-chunk piece logic); the halo chunks around them through
    // the generic loader.  c: the chunk whose MFMAs run next; chunk c + 1 loads.
    const int cm0 = (td.wlen + kGK - 1) / kGK, cm1 = max(cm0, (td.wlen + td.n_main) / kGK);
    int c = 0;
    auto step = [&](auto load_next) {
        const int buf = c & 1;
        load_next((c + 1) * kGK);
        // D of chunk c + 1 into the other buffer (read last before the previous barrier)
        gemm_dma_d(dr, dlds + (buf ^ 1) * kGemmDChunk, c + 1);
        if (active) gemm_chunk<E, DIR>(ulds, dlds, buf, lg, og, acc);
        gemm_stage<E, DIR, HASB>(ulds, buf ^ 1, u, sc, dot, gv);
        __syncthreads();  // also waits for the D copy
        ++c;
    };
    auto generic = [&](int k0) { gemm_load_u<DIR, HASB>(r, td, L, line_ok, k0, dotr, dot, u); };
    while (c + 1 < min(cm0, nch)) step(generic);
    {
        const MainU m = main_u<DIR, HASB>(r, td, line_ok);
        // the same step with the LDS buffer parity known at compile time
        // (LDS addresses as immediate offsets): two chunks per iteration
        auto step_p = [&](auto parity) {
            constexpr int buf = decltype(parity)::value;
            gemm_load_u_main<DIR, HASB>(m, (c + 1) * kGK, dotr, dot, u);
            gemm_dma_d(dr, dlds + (buf ^ 1) * kGemmDChunk, c + 1);
            if (active) gemm_chunk<E, DIR>(ulds, dlds, buf, lg, og, acc);
            gemm_stage<E, DIR, HASB>(ulds, buf ^ 1, u, sc, dot, gv);
            __syncthreads();
            ++c;
        };
        using P0 = std::integral_constant<int, 0>;
        using P1 = std::integral_constant<int, 1>;
        if ((c & 1) && c + 1 < cm1) step_p(P1{});
        while (c + 2 < cm1) {
            step_p(P0{});
            step_p(P1{});
        }
        if (c + 1 < cm1) step_p(P0{});
    }
    while (c + 1 < nch) step(generic);
    if (active) gemm_chunk<E, DIR>(ulds, dlds, c & 1, lg, og, acc);
    __syncthreads();
    // epilogue: L = C / (2^e 2^d_exp); out = t_n - s*coef*L (Solver::update_field_values)
    const float sc_upd = a.s * td.coef, inv = ldexpf(1.f, -(e + gt.d_exp));
    const bool use_pml = a.last && td.is_pml;
    const Rsrc rt = block_rsrc(a.tn, td.m_dom, td.m_size);
    const Rsrc rp = block_rsrc(a.pml, td.m_dom, td.m_size);
    const Rsrc ro = block_rsrc(a.out, td.m_dom, td.m_size);
    float mx = 0.f;
    if (active) {
        // value q of block (i, j): row 32 i + 8 (q >> 2) + 4 (lane >> 5) + (q & 3), column 32 j + (lane & 31)
        // of the wave's 64 x 64; all t_n loads are issued before the first is used
        const int lrow0 = DIR == 0 ? 64 * lg + 4 * (lane >> 5) : 64 * lg + (lane & 31);
        const int nrow0 = gt.n0 + (DIR == 0 ? 64 * og + (lane & 31) : 64 * og + 4 * (lane >> 5));
        auto offset = [&](int i, int j, int q, bool &ok) {
            const int rowd = 32 * i + 8 * (q >> 2) + (q & 3), cold = 32 * j;
            const int line = lrow0 + (DIR == 0 ? rowd : cold), n = nrow0 + (DIR == 0 ? cold : rowd);
            ok = line < td.count && n < gt.n_g;
            return ok ? (td.o_f + (td.line0 + line) * td.o_ls + n * td.o_ps) * 4 : kOOB;
        };
        float tnv[2][2][16];
        if (64 * lg + 64 <= td.count && 64 * og + 64 <= gt.n_g - gt.n0) {
            // the wave's whole 64 x 64 block lies inside the tile: a per-lane
            // offset and, per value, a uniform one (the buffer instruction's
            // scalar offset; no address arithmetic on the vector unit)
            const int base = (td.o_f + (td.line0 + lrow0) * td.o_ls + nrow0 * td.o_ps) * 4;
            const int sl = td.o_ls * 4, sp = td.o_ps * 4;
            auto uoff = [&](int i, int j, int q) {
                const int rowd = 32 * i + 8 * (q >> 2) + (q & 3), cold = 32 * j;
                return DIR == 0 ? rowd * sl + cold * sp : cold * sl + rowd * sp;
            };
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int q = 0; q < 16; ++q)
                        tnv[i][j][q] = __builtin_bit_cast(
                            float, __builtin_amdgcn_raw_buffer_load_b32(rt, base, uoff(i, j, q), 0));
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int q = 0; q < 16; ++q) {
                        const int so = uoff(i, j, q);
                        const float nv = __builtin_fmaf(-sc_upd, acc[i][j][q] * inv, tnv[i][j][q]);
                        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, nv), ro, base, so, 0);
                        if (a.last) {
                            const float pm = use_pml ? __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                                                  rp, base, so, 0))
                                                     : 1.f;
                            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, nv * pm), rt, base, so, 0);
                        }
                        mx = fmaxf(mx, fabsf(nv));
                    }
        } else {
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int q = 0; q < 16; ++q) {
                        bool ok;
                        tnv[i][j][q] = bload(rt, offset(i, j, q, ok));
                    }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int q = 0; q < 16; ++q) {
                        bool ok;
                        const int off = offset(i, j, q, ok);
                        const float nv = __builtin_fmaf(-sc_upd, acc[i][j][q] * inv, tnv[i][j][q]);
                        bstore(ro, off, nv);
                        if (a.last) bstore(rt, off, nv * (use_pml ? bload(rp, off) : 1.f));
                        if (ok) mx = fmaxf(mx, fabsf(nv));
                    }
        }
    }
    if (dot) {
        // output n_out - 1 of every line: the FP32 dot products (no scaling)
        const int n = gt.n_g;
        auto write_dot = [&](int line, float Ldot) {
            if (line < 0 || line >= td.count) return;
            const int off = (td.o_f + (td.line0 + line) * td.o_ls + n * td.o_ps) * 4;
            const float tnv = bload(rt, off), pm = use_pml ? bload(rp, off) : 1.f;
            const float nv = __builtin_fmaf(-sc_upd, Ldot, tnv);
            bstore(ro, off, nv);
            if (a.last) bstore(rt, off, nv * pm);
            mx = fmaxf(mx, fabsf(nv));
        };
        if constexpr (DIR == 0) {
            // the 4 threads of a line are adjacent lanes; slot q holds line tid/4 + 64 q
#pragma unroll
            for (int q = 0; q < kGSlots; ++q) {
                float g = gv[q];
                g += __shfl_xor(g, 1);
                g += __shfl_xor(g, 2);
                write_dot((tid & 3) == 0 ? stage_line<DIR>(tid, q) : -1, g);
            }
        } else {
            // partial sums of the 4 k-groups (2 per thread) through LDS (the ring is no longer read)
            float *red = reinterpret_cast<float *>(gemm_lds);
            red[(tid >> 7) * kGL + (tid & (kGL - 1))] = gv[0] + gv[1];
            __syncthreads();
            float Ldot = 0.f;
            if (tid < kGL) {
#pragma unroll
                for (int k = 0; k < kGemmThreads / kGL; ++k) Ldot += red[k * kGL + tid];
            }
            write_dot(tid < kGL ? tid : -1, Ldot);
        }
    }
    track_max(a, td.scene, td.vel ? 0 : 1, mx);
}

template <typename E>
__global__ void __launch_bounds__(kGemmThreads) __attribute__((amdgpu_waves_per_eu(2))) stage_gemm_kernel(StageArgs a, const GemmTile *__restrict__ tiles,
                                                                  const float *__restrict__ ftab) {
    // clears the slot the next stage writes (nothing reads or writes it in this stage)
    if (blockIdx.x == 0 && a.umax)
        for (int i = threadIdx.x; i < 2 * a.n_scenes; i += blockDim.x) a.umax[(size_t)a.slot_z * a.n_scenes * 2 + i] = 0u;
    extern __shared__ __attribute__((aligned(16))) unsigned char gemm_lds[];
    // XCD-contiguous runs of tiles always (the tiles of one matrix and of one
    // line block share an XCD's L2), whatever order the FFT-form tiles use
    const GemmTile gt = tiles[tile_index(blockIdx.x, gridDim.x, 1)];
    const bool hasB = !gt.t.vel && a.p0b != nullptr;  // p0 = px + py
    if (gt.t.dir == 0) {
        if (hasB) gemm_tile<E, 0, true>(a, gt, ftab, gemm_lds);
        else gemm_tile<E, 0, false>(a, gt, ftab, gemm_lds);
    } else {
        if (hasB) gemm_tile<E, 1, true>(a, gt, ftab, gemm_lds);
        else gemm_tile<E, 1, false>(a, gt, ftab, gemm_lds);
    }
}

// Largest |p0| and |vx|, |vy| of every scene in the state a first stage reads
// after the host wrote fields (it has no previous stage): into slot `slot`.
struct MaxRange {
    long long off;  // element offset from the arena base
    long long n;
    int idx;        // scene * 2 + class
};

__global__ void fill_zero_kernel(unsigned *__restrict__ p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

__global__ void state_max_kernel(const float *__restrict__ arena, unsigned *__restrict__ umax,
                                 const MaxRange *__restrict__ ranges, int nr, int n_scenes, int slot) {
    for (int k = blockIdx.y; k < nr; k += gridDim.y) {
        const MaxRange r = ranges[k];
        float m = 0.f;
        for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < r.n; e += (long long)gridDim.x * blockDim.x)
            m = fmaxf(m, fabsf(arena[r.off + e]));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
        if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax(&umax[(size_t)slot * n_scenes * 2 + r.idx], __float_as_uint(m));
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

// Time steps per captured hipGraph unless pstd_ctx_set_graph_steps says otherwise.
constexpr int kDefaultGraphSteps = 20;

struct pstd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // with `overlap`, GEMM-form stages run their FFT-form tiles on `side`,
    // concurrently with the GEMM kernel on `stream` (fork and join through the
    // two events; both streams at the default priority).  Off by default: the
    // GEMM kernel fills every CU's LDS, so the two only trade places (measured
    // +1 % on 64 scenes, 0 on 256 sources, -4 % on one 4096^2 domain)
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool overlap = false;
    std::string err;
    bool planned = false;
    // plan
    std::vector<DomainDev> doms;
    std::vector<pstd_plan_domain> plan_domains;
    std::vector<TileDesc> tiles_small, tiles_4096, tiles_8192;
    // FFT-form segments (line0/count = the segment's line range) and the
    // passthrough/receiver tiles: the tile lists are cut from them for the
    // block size of the F <= 2048 class (small_nt, autotuned: 256 or 512)
    std::vector<TileDesc> fft_segs, tiles_extra, tiles_all;
    int small_nt = 256;
    int small_fmax = 0;  // largest F of the F <= 2048 class (128 threads need F <= 1024)
    std::vector<GemmTile> tiles_gemm;
    int n_scenes = 1;
    std::vector<MaxRange> max_ranges;  // state read by a first stage after host writes (GEMM form)
    int form = PSTD_FORM_FFT;      // the form the uploaded plan runs in
    int form_req = PSTD_FORM_FFT;  // as requested (PSTD_FORM_AUTO: chosen at upload)
    double gemm_flops = 0;   // algorithmic MFMA flops of one stage (3 split products)
    long long gemm_cells = 0;
    size_t arena_floats = 0;
    float dt = 0, c1 = 0;
    float alpha[6] = {0};
    int nrec = 0;
    // device buffers
    float *arena = nullptr;  // 4 arenas back to back: TN, S0, S1, P0I
    float *pml = nullptr;
    float2 *ctab = nullptr;
    float *ftab = nullptr;
    TileDesc *d_tiles = nullptr;  // small | 4096 | 8192
    GemmTile *d_gemm = nullptr;
    void *dmat = nullptr;         // derivative matrices of the GEMM form (hi, lo parts; 2-byte elements)
    unsigned *umax = nullptr;     // per-scene maxima of the written values, 3 slots x scenes x 2 classes
    MaxRange *d_max_ranges = nullptr;
    ReceiverCell *d_cells = nullptr;
    // spectral receivers: line jobs, per-receiver jobs, weights, column buffer
    SILine *d_si_lines = nullptr;
    SIPoint *d_si_points = nullptr;
    float *d_si_coef = nullptr;
    float *d_si_col = nullptr;
    int n_si_lines = 0, n_si = 0;
    P0Range *d_ranges = nullptr;
    int n_ranges = 0;
    float *d_samples = nullptr;
    int sample_cap = 0;
    int steps_done = 0;
    bool p0_explicit = true;
    int xcd_remap = 1;
    int lean = 0;  // stage kernel flavour (kEarly, kLate, kMid)
    float tuned_stage_ms = 0;
    // the GEMM form's maxima in `umax` no longer describe the state (an
    // allocation-free advance used the shared array): recompute before the next step
    bool max_stale = false;
    // per-kernel-class timing (pstd_ctx_profile_kernels): event pairs around
    // every launch of a class while `prof` is set
    bool prof = false;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof_events;
    // hipGraph of `graph_steps` whole time steps (0: launches are issued one by
    // one).  The graph writes its receiver samples into `d_gsamples`; each replay
    // is followed by a device-side copy to its place in `d_samples`.
    int graph_steps = kDefaultGraphSteps;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    float *d_gsamples = nullptr;
    // autotuner bookkeeping: time steps its trials ran, and whether the launch
    // shape came from the per-(device, plan) cache instead
    int trial_steps = 0;
    bool tune_cached = false;

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (d_gsamples) (void)hipFree(d_gsamples);
        graph_exec = nullptr;
        graph = nullptr;
        d_gsamples = nullptr;
    }

    float *TN() const { return arena; }
    float *S(int i) const { return arena + (1 + i) * arena_floats; }
    float *P0I() const { return arena + 3 * arena_floats; }

    void release() {
        drop_graph();
        (void)hipFree(arena);
        (void)hipFree(pml);
        (void)hipFree(ctab);
        (void)hipFree(ftab);
        (void)hipFree(d_tiles);
        (void)hipFree(d_gemm);
        (void)hipFree(dmat);
        (void)hipFree(umax);
        (void)hipFree(d_max_ranges);
        d_gemm = nullptr;
        dmat = nullptr;
        umax = nullptr;
        d_max_ranges = nullptr;
        (void)hipFree(d_cells);
        (void)hipFree(d_si_lines);
        (void)hipFree(d_si_points);
        (void)hipFree(d_si_coef);
        (void)hipFree(d_si_col);
        d_si_lines = nullptr;
        d_si_points = nullptr;
        d_si_coef = d_si_col = nullptr;
        n_si_lines = n_si = 0;
        (void)hipFree(d_ranges);
        (void)hipFree(d_samples);
        arena = pml = ftab = d_samples = nullptr;
        ctab = nullptr;
        d_tiles = nullptr;
        d_cells = nullptr;
        d_ranges = nullptr;
        planned = false;
    }
};

namespace {

// Device plan cache entries: the tables of plan_tables.h rounded once to float.
void build_filter(int F, int L, int field, double dx, std::vector<float2> &out) {
    const std::vector<std::complex<double>> h = pstdamd::derivative_filter(F, L, field, dx);
    out.resize(F);
    for (int m = 0; m < F; ++m) out[m] = make_float2((float)h[m].real(), (float)h[m].imag());
}

void build_twiddles(int F, std::vector<float2> &out) {
    out.resize(F);
    for (int i = 0; i < F; ++i) {
        double a = -2.0 * M_PI * (double)i / (double)F;
        out[i] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
}

void build_window(int w, int patch_error, std::vector<float> &out) {
    const std::vector<double> win = pstdamd::window_coefficients(w, patch_error);
    out.assign(win.begin(), win.end());
}

// Kernel classes of pstd_ctx_profile_kernels.
constexpr int kClassSmall = 0, kClass4096 = 1, kClass8192 = 2, kClassGemm = 3, kClasses = 4;

// Brackets the launches issued by `launch` with an event pair when profiling.
template <class Fn>
void timed_launch(pstd_ctx *ctx, hipStream_t st, int cls, Fn launch) {
    if (!ctx->prof) {
        launch();
        return;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        launch();
        return;
    }
    (void)hipEventRecord(e0, st);
    launch();
    (void)hipEventRecord(e1, st);
    ctx->prof_events.push_back({cls, {e0, e1}});
}

// One RK stage.  sample_slot >= 0 (first stage only): the launch also records
// the receivers of the previous step into that sample slot.
int launch_stage(pstd_ctx *ctx, int k, int sample_slot) {
    StageArgs a{};
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
    a.cells = ctx->d_cells;
    a.nrec = ctx->nrec;
    a.samples = ctx->d_samples;
    a.slot = (k == 0 && !ctx->p0_explicit && ctx->nrec > 0) ? sample_slot : -1;
    a.tn = ctx->TN();
    a.out = ctx->S(outi);
    a.pml = ctx->pml;
    a.s = ctx->dt * ctx->alpha[k];
    a.last = k == 5;
    a.remap = ctx->xcd_remap;
    const int ng = (int)ctx->tiles_gemm.size();
    if (ng) {
        // stage k writes slot k%3, reads (k+2)%3 and clears (k+1)%3 (6 stages per
        // step: the same slots in every step, so a captured graph replays them)
        a.umax = ctx->umax;
        a.slot_w = k % 3;
        a.slot_r = (k + 2) % 3;
        a.slot_z = (k + 1) % 3;
        a.n_scenes = ctx->n_scenes;
        a.dmat = ctx->dmat;
        if (k == 0 && ctx->p0_explicit && !ctx->max_ranges.empty()) {
            const int nr = (int)ctx->max_ranges.size();
            hipLaunchKernelGGL(state_max_kernel, dim3(16, std::min(nr, 4096)), dim3(256), 0, ctx->stream, ctx->arena,
                               ctx->umax, ctx->d_max_ranges, nr, ctx->n_scenes, a.slot_r);
        }
    }
    const int ns = (int)ctx->tiles_small.size(), n4 = (int)ctx->tiles_4096.size(), n8 = (int)ctx->tiles_8192.size();
    const int fl = ctx->lean;
    // FFT-form tiles beside a GEMM launch: on the side stream, forked after
    // everything issued so far and joined before anything issued later
    const bool fork = ng && (ns || n4 || n8) && ctx->overlap && ctx->side;
    const hipStream_t fs = fork ? ctx->side : ctx->stream;
    if (fork) {
        HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream));
        HIP_TRY(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
    }
    if (ns) timed_launch(ctx, fs, kClassSmall, [&] {
        a.tiles = ctx->d_tiles;
        if (ctx->small_nt == 512) {
            if (fl == kLate) hipLaunchKernelGGL((stage_small_kernel<512, kLate>), dim3(ns), dim3(512), 0, fs, a);
            else if (fl == kMid) hipLaunchKernelGGL((stage_small_kernel<512, kMid>), dim3(ns), dim3(512), 0, fs, a);
            else hipLaunchKernelGGL((stage_small_kernel<512, kEarly>), dim3(ns), dim3(512), 0, fs, a);
        } else if (ctx->small_nt == 128) {
            if (fl == kLate) hipLaunchKernelGGL((stage_small_kernel<128, kLate>), dim3(ns), dim3(128), 0, fs, a);
            else if (fl == kMid) hipLaunchKernelGGL((stage_small_kernel<128, kMid>), dim3(ns), dim3(128), 0, fs, a);
            else hipLaunchKernelGGL((stage_small_kernel<128, kEarly>), dim3(ns), dim3(128), 0, fs, a);
        } else {
            if (fl == kLate) hipLaunchKernelGGL((stage_small_kernel<256, kLate>), dim3(ns), dim3(256), 0, fs, a);
            else if (fl == kMid) hipLaunchKernelGGL((stage_small_kernel<256, kMid>), dim3(ns), dim3(256), 0, fs, a);
            else hipLaunchKernelGGL((stage_small_kernel<256, kEarly>), dim3(ns), dim3(256), 0, fs, a);
        }
    });
    if (n4) timed_launch(ctx, fs, kClass4096, [&] {
        a.tiles = ctx->d_tiles + ns;
        size_t lds = 2 * sizeof(float2) * LdsMap<4096, 1, 0>::SIZE;
        if (fl == kLate) hipLaunchKernelGGL((stage_big_kernel<4096, kLate>), dim3(n4), dim3(512), lds, fs, a);
        else if (fl == kMid) hipLaunchKernelGGL((stage_big_kernel<4096, kMid>), dim3(n4), dim3(512), lds, fs, a);
        else hipLaunchKernelGGL((stage_big_kernel<4096, kEarly>), dim3(n4), dim3(512), lds, fs, a);
    });
    if (n8) timed_launch(ctx, fs, kClass8192, [&] {
        // one line pair per 512-thread block as two 4096-point transforms
        a.tiles = ctx->d_tiles + ns + n4;
        hipLaunchKernelGGL(stage_split_kernel, dim3(n8), dim3(kSplitT), kSplitLds, fs, a);
    });
    if (fork) HIP_TRY(hipEventRecord(ctx->ev_join, ctx->side));
    if (ng) timed_launch(ctx, ctx->stream, kClassGemm, [&] {
        if (ctx->form == PSTD_FORM_GEMM_BF16)
            hipLaunchKernelGGL(stage_gemm_kernel<__bf16>, dim3(ng), dim3(kGemmThreads), kGemmLdsBytes, ctx->stream, a,
                               ctx->d_gemm, ctx->ftab);
        else
            hipLaunchKernelGGL(stage_gemm_kernel<_Float16>, dim3(ng), dim3(kGemmThreads), kGemmLdsBytes, ctx->stream, a,
                               ctx->d_gemm, ctx->ftab);
    });
    if (fork) HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
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

// arena element offset of a field (p0 lives in the px slot of P0I)
long long field_off(const DomainDev &d, int f) {
    switch (f) {
        case PSTD_FIELD_VX: return d.base + d.f_vx;
        case PSTD_FIELD_VY: return d.base + d.f_vy;
        case PSTD_FIELD_PY: return d.base + d.f_py;
        default: return d.base + d.f_px;
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

// Device address of `scratch` per device (the array's name is a host-side
// stand-in: kernels get this address).  Resolved once per device with that
// device current.
namespace {
std::mutex g_scratch_mu;
std::map<int, unsigned *> g_scratch;

int resolve_scratch(int device, unsigned **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch.find(device);
    if (it == g_scratch.end()) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return PSTD_ERR_DEVICE;
        it = g_scratch.emplace(device, static_cast<unsigned *>(ptr)).first;
    }
    *out = it->second;
    return PSTD_OK;
}
}  // namespace

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
    // the side stream and its events (pstd_ctx_set_stage_overlap refuses to
    // turn the overlap on if one of them could not be created)
    if (hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess) ctx->side = nullptr;
    if (hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess) ctx->ev_fork = nullptr;
    if (hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess) ctx->ev_join = nullptr;
    // the long-line classes need more than 64 KiB of LDS per workgroup
    const size_t lds4 = 2 * sizeof(float2) * LdsMap<4096, 1, 0>::SIZE;
    const hipFuncAttribute attr = hipFuncAttributeMaxDynamicSharedMemorySize;
    const void *big4[3] = {(const void *)stage_big_kernel<4096, kEarly>, (const void *)stage_big_kernel<4096, kLate>,
                           (const void *)stage_big_kernel<4096, kMid>};
    bool lds_ok = hipFuncSetAttribute((const void *)stage_split_kernel, attr, kSplitLds) == hipSuccess;
    for (int f = 0; f < 3; ++f) lds_ok = lds_ok && hipFuncSetAttribute(big4[f], attr, lds4) == hipSuccess;
    lds_ok = lds_ok &&
             hipFuncSetAttribute((const void *)stage_gemm_kernel<_Float16>, attr, kGemmLdsBytes) == hipSuccess &&
             hipFuncSetAttribute((const void *)stage_gemm_kernel<__bf16>, attr, kGemmLdsBytes) == hipSuccess;
    if (!lds_ok) {
        ctx->err = "cannot reserve LDS for the long-line kernels";
    }
    unsigned *sp = nullptr;
    if (resolve_scratch(device, &sp) != PSTD_OK) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return PSTD_ERR_DEVICE;
    }
    *out = ctx;
    return PSTD_OK;
}

void pstd_ctx_destroy(pstd_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->release();
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *pstd_ctx_last_error(const pstd_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// Enqueues `chunk` time steps: 6 stage launches per step (the first stage of
// step s > 0 also samples step s-1 into sample slot base+s-1) and the sampling
// of the chunk's last step into slot base+chunk-1.
static int enqueue_steps(pstd_ctx *ctx, int chunk, int base = 0) {
    for (int s = 0; s < chunk; ++s) {
        for (int k = 0; k < 6; ++k) {
            int rc = launch_stage(ctx, k, s == 0 ? -1 : base + s - 1);
            if (rc) return rc;
            if (k == 0) ctx->p0_explicit = false;
        }
        if (ctx->n_si > 0) {
            // spectral receivers of this step, from its raw stage-5 output
            hipLaunchKernelGGL(si_line_kernel, dim3((ctx->n_si_lines + 3) / 4), dim3(256), 0, ctx->stream, ctx->S(1),
                               ctx->d_si_lines, ctx->n_si_lines, ctx->d_si_coef, ctx->d_si_col);
            hipLaunchKernelGGL(si_point_kernel, dim3(ctx->n_si), dim3(64), 0, ctx->stream, ctx->d_si_col,
                               ctx->d_si_points, ctx->d_si_coef, ctx->d_samples, ctx->nrec, base + s);
            HIP_TRY(hipGetLastError());
        }
    }
    if (ctx->nrec > 0) {
        hipLaunchKernelGGL(receiver_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->S(1), ctx->d_cells, ctx->nrec,
                           ctx->d_samples, base + chunk - 1);
        HIP_TRY(hipGetLastError());
    }
    return PSTD_OK;
}

// Tile lists for the current block size of the F <= 2048 class: a segment's
// lines in tiles of NL = 2 P lines (P = NT / (F/8) pairs; one pair per tile
// above 2048), then the passthrough/receiver tiles; uploaded as
// small | 4096 | 8192.  The pairs (line0 + 2p, line0 + 2p + 1) are the same
// for every NT, so results do not depend on it.
static int upload_tiles(pstd_ctx *ctx) {
    ctx->tiles_small.clear();
    ctx->tiles_4096.clear();
    ctx->tiles_8192.clear();
    for (const TileDesc &g : ctx->fft_segs) {
        const int per = 2 * (g.F >= 4096 ? 1 : ctx->small_nt / (g.F / 8));
        for (int l = 0; l < g.count; l += per) {
            TileDesc t = g;
            t.line0 = g.line0 + l;
            t.count = std::min(per, g.count - l);
            if (g.F == 4096) ctx->tiles_4096.push_back(t);
            else if (g.F == 8192) ctx->tiles_8192.push_back(t);
            else ctx->tiles_small.push_back(t);
        }
    }
    ctx->tiles_small.insert(ctx->tiles_small.end(), ctx->tiles_extra.begin(), ctx->tiles_extra.end());
    ctx->tiles_all = ctx->tiles_small;
    ctx->tiles_all.insert(ctx->tiles_all.end(), ctx->tiles_4096.begin(), ctx->tiles_4096.end());
    ctx->tiles_all.insert(ctx->tiles_all.end(), ctx->tiles_8192.begin(), ctx->tiles_8192.end());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->d_tiles) HIP_TRY(hipFree(ctx->d_tiles));
    ctx->d_tiles = nullptr;
    HIP_TRY(hipMalloc(&ctx->d_tiles, std::max<size_t>(1, ctx->tiles_all.size()) * sizeof(TileDesc)));
    if (!ctx->tiles_all.empty())
        HIP_TRY(hipMemcpyAsync(ctx->d_tiles, ctx->tiles_all.data(), ctx->tiles_all.size() * sizeof(TileDesc),
                               hipMemcpyHostToDevice, ctx->stream));
    return PSTD_OK;
}

// Launch-shape autotuner.  Times every (block size, stage-kernel flavour, tile
// order) candidate on the freshly uploaded (all-zero) state and keeps the
// fastest.  Zero fields stay exactly zero under the update, and the arenas are
// cleared again afterwards, so tuning leaves no trace in the results.
// The autotuner's choices, per (device, plan signature): a later context with a
// plan of the same launch geometry on the same device takes the choice and
// runs no trials.
struct TuneChoice {
    int lean, xcd_remap, small_nt;
    float stage_ms;
};
std::mutex g_tune_mu;
std::map<std::pair<int, uint64_t>, TuneChoice> g_tune;
// PSTD_FORM_AUTO's choice per (device, signature of the plan's FFT-form upload)
std::map<std::pair<int, uint64_t>, int> g_form;

// FNV-1a over what the launches' timing depends on: the form, the scene count,
// every tile's line class, line count and direction, and the arena size.  Two
// plans that hash alike launch the same grids of the same kernels.
uint64_t plan_signature(const pstd_ctx *ctx) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](long long v) {
        for (int i = 0; i < 8; ++i) {
            h ^= (uint64_t)((v >> (8 * i)) & 0xff);
            h *= 1099511628211ull;
        }
    };
    mix(ctx->form);
    mix(ctx->n_scenes);
    mix((long long)ctx->arena_floats);
    mix((long long)ctx->fft_segs.size());
    for (const TileDesc &t : ctx->fft_segs) {
        mix(t.kind);
        mix(t.F);
        mix(t.count);
        mix(t.dir);
        mix(t.vel);
        mix(t.n_out);
    }
    mix((long long)ctx->tiles_extra.size());
    mix((long long)ctx->tiles_gemm.size());
    for (const GemmTile &g : ctx->tiles_gemm) {
        mix(g.t.dir);
        mix(g.t.vel);
        mix(g.t.count);
        mix(g.n_rows);
        mix(g.k_pad);
    }
    return h;
}

static int autotune(pstd_ctx *ctx) {
    const std::pair<int, uint64_t> key(ctx->device, plan_signature(ctx));
    ctx->trial_steps = 0;
    ctx->tune_cached = false;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            ctx->lean = it->second.lean;
            ctx->xcd_remap = it->second.xcd_remap;
            ctx->tuned_stage_ms = it->second.stage_ms;
            ctx->tune_cached = true;
            if (it->second.small_nt != ctx->small_nt) {
                ctx->small_nt = it->second.small_nt;
                const int rc = upload_tiles(ctx);
                if (rc) return rc;
            }
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            return PSTD_OK;
        }
    }
    static const int cand[6][2] = {{kLate, 1}, {kMid, 1}, {kEarly, 1}, {kLate, 0}, {kMid, 0}, {kEarly, 0}};  // (flavour, xcd order)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    // one warm-up step, then GPU ms of n steps
    int n = 3;
    auto time_steps = [&](float &ms) -> int {
        ctx->trial_steps += 1 + n;
        int rc = enqueue_steps(ctx, 1);
        if (rc == PSTD_OK && hipEventRecord(e0, ctx->stream) != hipSuccess) rc = PSTD_ERR_DEVICE;
        if (rc == PSTD_OK) rc = enqueue_steps(ctx, n);
        if (rc == PSTD_OK && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = PSTD_ERR_DEVICE;
        if (rc == PSTD_OK && hipEventSynchronize(e1) != hipSuccess) rc = PSTD_ERR_DEVICE;
        if (rc == PSTD_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = PSTD_ERR_DEVICE;
        return rc;
    };
    float best = 0;
    int bi = 0, bnt = ctx->small_nt, rc = PSTD_OK;
    bool first = true;
    {
        // small (latency-bound) scenes get more timed steps per candidate
        // (about 2 ms of GPU time each), so that run-to-run noise does not
        // pick the shape; large ones keep 3
        float ms = 0;
        rc = time_steps(ms);
        if (rc == PSTD_OK && ms > 0.f) n = std::min(50, std::max(3, (int)std::ceil(2.0f / (ms / 3.0f))));
    }
    // block sizes of the F <= 2048 class (it alone depends on them)
    std::vector<int> nts = {ctx->small_nt};
    if (ctx->small_fmax > 0)
        for (int nt : {128, 256, 512})
            if (nt != ctx->small_nt && (nt >= 256 || ctx->small_fmax <= 1024)) nts.push_back(nt);
    for (int nt : nts) {
        if (rc != PSTD_OK) break;
        if (nt != ctx->small_nt) {
            ctx->small_nt = nt;
            rc = upload_tiles(ctx);
        }
        for (int c = 0; c < 6 && rc == PSTD_OK; ++c) {
            ctx->lean = cand[c][0];
            ctx->xcd_remap = cand[c][1];
            float ms = 0;
            rc = time_steps(ms);
            if (rc == PSTD_OK && (first || ms < best)) {
                best = ms;
                bi = c;
                bnt = nt;
                first = false;
            }
        }
    }
    if (rc == PSTD_OK) {
        ctx->lean = cand[bi][0];
        ctx->xcd_remap = cand[bi][1];
        if (bnt != ctx->small_nt) {
            ctx->small_nt = bnt;
            rc = upload_tiles(ctx);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        if (ctx->err.empty()) ctx->err = "autotune: HIP event failure";
        return rc;
    }
    ctx->tuned_stage_ms = best / (6.0 * n);
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tune[key] = {ctx->lean, ctx->xcd_remap, ctx->small_nt, ctx->tuned_stage_ms};
    }
    HIP_TRY(hipMemsetAsync(ctx->arena, 0, 4 * (size_t)ctx->arena_floats * sizeof(float), ctx->stream));
    if (ctx->umax) HIP_TRY(hipMemsetAsync(ctx->umax, 0, 3 * 2 * (size_t)ctx->n_scenes * sizeof(unsigned), ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->steps_done = 0;
    ctx->p0_explicit = true;
    return PSTD_OK;
}

// 16-bit parts of the GEMM form, from double, rounded to nearest even.
static uint16_t to_f16_bits(double v) {
    const _Float16 h = (_Float16)v;
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}
static double from_f16_bits(uint16_t b) {
    _Float16 h;
    std::memcpy(&h, &b, 2);
    return (double)h;
}
static uint16_t to_bf16_bits(double v) {
    if (v == 0.0) return 0;
    int e = 0;
    const double m = std::frexp(v, &e);                        // |m| in [0.5, 1)
    const double q = std::nearbyint(std::ldexp(m, 8));         // 8 significant bits, ties to even
    const float f = (float)std::ldexp(q, e - 8);               // exact in float
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (uint16_t)(u >> 16);
}
static double from_bf16_bits(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return (double)f;
}

// Whether a segment runs in the GEMM form: its lines and outputs fill blocks of
// 128 lines x 256 outputs at least to 3/4 (thin PML strips keep the FFT form).
static bool gemm_fills(int lines, int n_g) {
    constexpr int kFillN = 256;
    const double cap = (double)((lines + kGL - 1) / kGL * kGL) * (double)((n_g + kFillN - 1) / kFillN * kFillN);
    return 4.0 * lines * (double)n_g >= 3.0 * cap;
}

// GEMM form of one segment: its derivative matrix (built once per distinct
// segment shape and stencil, split into 16-bit hi + lo parts of D * 2^d_exp;
// for pressure lines the last row kept apart in FP32 for the dot product) and
// its 128-line x 128-output tiles.
static int add_gemm_tiles(pstd_ctx *ctx, const pstd_plan_desc *plan, const pstd_plan_segment &s, const TileDesc &g,
                          std::vector<uint16_t> &dmat, std::vector<float> &ftab,
                          std::map<std::string, std::tuple<long long, int, int>> &dcache) {
    pstdamd::SegmentShape sh{g.F, s.table_len, s.field, (double)plan->dx, s.wlen, g.n_main, (double)plan->patch_error,
                             (double)s.c_side1, (double)s.c_main1, (double)s.c_side2, (double)s.c_main2};
    const int K = sh.K(), n_out = sh.n_out();
    const bool dot_row = s.field == 0;  // n + 1 outputs: the last one as an FP32 dot product
    const int n_g = n_out - (dot_row ? 1 : 0);
    const int k_pad = (K + kGK - 1) / kGK * kGK, n_rows = (n_g + kGN - 1) / kGN * kGN, nch = k_pad / kGK;
    const bool bf16 = ctx->form == PSTD_FORM_GEMM_BF16;
    char key[256];
    std::snprintf(key, sizeof key, "%d/%d/%d/%d/%d/%a/%a/%a/%a", sh.F, sh.table_len, sh.field, sh.w, sh.n_main,
                  sh.c_s1, sh.c_m1, sh.c_s2, sh.c_m2);
    auto it = dcache.find(key);
    if (it == dcache.end()) {
        std::vector<double> D;
        try {
            D = pstdamd::derivative_matrix(sh);
        } catch (const std::exception &e) {
            ctx->err = std::string("GEMM form: ") + e.what();
            return PSTD_ERR_INVALID;
        }
        double dm = 0;
        for (int i = 0; i < n_g; ++i)
            for (int k = 0; k < K; ++k) dm = std::max(dm, std::fabs(D[(size_t)i * K + k]));
        const int d_exp = dm > 0 ? 14 - std::ilogb(dm) : 0;
        // LDS images: for every 128-output block b and chunk c, [hi, lo][128 rows][kGP]
        const long long hi = (long long)dmat.size();
        dmat.resize(dmat.size() + (size_t)(n_rows / kGN) * nch * kGemmDChunk, 0);
        for (int i = 0; i < n_g; ++i)
            for (int k = 0; k < K; ++k) {
                const double v = std::ldexp(D[(size_t)i * K + k], d_exp);
                const uint16_t h = bf16 ? to_bf16_bits(v) : to_f16_bits(v);
                const double hv = bf16 ? from_bf16_bits(h) : from_f16_bits(h);
                const size_t at = hi + ((size_t)(i / kGN) * nch + k / kGK) * kGemmDChunk + (size_t)(i % kGN) * kGP + k % kGK;
                dmat[at] = h;
                dmat[at + (size_t)kGN * kGP] = bf16 ? to_bf16_bits(v - hv) : to_f16_bits(v - hv);
            }
        int dot_off = -1;
        if (dot_row) {
            dot_off = (int)ftab.size();
            for (int k = 0; k < K; ++k) ftab.push_back((float)D[(size_t)(n_out - 1) * K + k]);
        }
        it = dcache.emplace(key, std::make_tuple(hi, d_exp, dot_off)).first;
    }
    const long long d_base = std::get<0>(it->second), d_len = (long long)nch * kGemmDChunk;
    ctx->gemm_flops += 3.0 * 2.0 * K * (double)n_out * (s.line_end - s.line_begin);
    ctx->gemm_cells += (long long)n_out * (s.line_end - s.line_begin);
    for (int l = s.line_begin; l < s.line_end; l += kGL) {
        GemmTile gt{};
        gt.t = g;
        gt.t.kind = 3;
        gt.t.line0 = l;
        gt.t.count = std::min(kGL, s.line_end - l);
        gt.d_len = d_len;
        gt.n_rows = n_rows;
        gt.k_pad = k_pad;
        gt.n_g = n_g;
        gt.d_exp = std::get<1>(it->second);
        gt.K = K;
        for (int n0 = 0; n0 < n_g; n0 += kGN) {
            gt.n0 = n0;
            gt.d_img = d_base + (long long)(n0 / kGN) * d_len;
            gt.dot_off = n0 + kGN >= n_g ? std::get<2>(it->second) : -1;  // the last output tile
            ctx->tiles_gemm.push_back(gt);
        }
    }
    return PSTD_OK;
}

// dom_scene: scene of every domain (a batch of scenes), or nullptr (one scene).
// Line jobs and weights of the spectral receivers (Receiver::compute_with_si,
// kernel/core/Receiver.cpp:85-121): for each, the Y pressure segment of its
// domain through its column gives the Y line (neighbours, stencil, window);
// every row of the domain and the w window rows of each Y neighbour are
// shifted along X with the X pressure segment holding that row; the weights
// are rows of the segments' line operators with the shift filter
// (plan_tables.h), rounded once to float.
static int build_spectral_receivers(pstd_ctx *ctx, const pstd_plan_desc *plan) {
    std::vector<SILine> lines;
    std::vector<SIPoint> points;
    std::vector<float> coef;
    int col_total = 0;
    std::map<std::pair<int, int>, int> row_cache;  // (segment, output column) -> weights, per receiver
    auto find_segment = [&](int dom, int dir, int line) -> int {
        for (int k = 0; k < plan->n_segments; ++k) {
            const pstd_plan_segment &s = plan->segments[k];
            if (s.domain == dom && s.direction == dir && s.field == 0 && s.line_begin <= line && line < s.line_end)
                return k;
        }
        return -1;
    };
    auto shape_of = [&](const pstd_plan_segment &s) {
        const DomainDev &d = ctx->doms[s.domain];
        const int n = s.direction == 0 ? d.nx : d.ny;
        return pstdamd::SegmentShape{next_pow2(n + 2 * s.wlen), s.table_len, 0, (double)plan->dx, s.wlen, n,
                                     (double)plan->patch_error, (double)s.c_side1, (double)s.c_main1,
                                     (double)s.c_side2, (double)s.c_main2};
    };
    auto add_weights = [&](const pstd_plan_segment &s, double offset, int out) {
        const pstdamd::SegmentShape sh = shape_of(s);
        const std::vector<double> row = pstdamd::operator_row(sh, pstdamd::shift_filter(sh.F, sh.table_len, offset), out);
        const int at = (int)coef.size();
        for (double v : row) coef.push_back((float)v);
        return at;
    };
    for (int r = 0; r < plan->n_receivers; ++r) {
        const pstd_plan_receiver &rc = plan->receivers[r];
        if (!rc.spectral) continue;
        if (!(rc.offset_x >= 0.f && rc.offset_x < 1.f && rc.offset_y >= 0.f && rc.offset_y < 1.f)) {
            ctx->err = "spectral receiver " + std::to_string(r) + ": offsets outside [0, 1)";
            return PSTD_ERR_INVALID;
        }
        const DomainDev &c = ctx->doms[rc.domain];
        const int cx = rc.index % c.nx, cy = rc.index / c.nx;
        const int ky = find_segment(rc.domain, 1, cx);
        if (ky < 0) {
            ctx->err = "spectral receiver " + std::to_string(r) + ": no Y pressure segment through its column";
            return PSTD_ERR_INVALID;
        }
        const pstd_plan_segment &sy = plan->segments[ky];
        const int w = sy.wlen, col0 = col_total;
        row_cache.clear();
        // the rows to shift along X: (domain, row, output column, column buffer element)
        struct Row {
            int dom, row, out, dst;
        };
        std::vector<Row> rows;
        for (int y = 0; y < c.ny; ++y) rows.push_back({rc.domain, y, cx, col0 + w + y});
        if (sy.side1 >= 0) {
            const DomainDev &n1 = ctx->doms[sy.side1];
            for (int j = 0; j < w; ++j) rows.push_back({sy.side1, n1.ny - w + j, cx + sy.side1_offset, col0 + j});
        }
        if (sy.side2 >= 0)
            for (int j = 0; j < w; ++j) rows.push_back({sy.side2, j, cx + sy.side2_offset, col0 + w + c.ny + j});
        for (const Row &q : rows) {
            const int kx = find_segment(q.dom, 0, q.row);
            if (kx < 0) {
                ctx->err = "spectral receiver " + std::to_string(r) + ": domain " + std::to_string(q.dom) +
                           " has no X pressure segment through row " + std::to_string(q.row);
                return PSTD_ERR_INVALID;
            }
            const pstd_plan_segment &sx = plan->segments[kx];
            const DomainDev &d = ctx->doms[q.dom];
            if (q.out < 0 || q.out >= d.nx) {
                ctx->err = "spectral receiver " + std::to_string(r) + ": column outside a neighbour";
                return PSTD_ERR_INVALID;
            }
            auto it = row_cache.find({kx, q.out});
            if (it == row_cache.end()) {
                int at;
                try {
                    at = add_weights(sx, (double)rc.offset_x, q.out);
                } catch (const std::exception &e) {
                    ctx->err = std::string("spectral receiver: ") + e.what();
                    return PSTD_ERR_INVALID;
                }
                it = row_cache.emplace(std::make_pair(kx, q.out), at).first;
            }
            SILine L{};
            L.m_px = d.base + d.f_px + (long long)q.row * d.nx;
            L.m_py = d.base + d.f_py + (long long)q.row * d.nx;
            L.s1_px = L.s1_py = L.s2_px = L.s2_py = -1;
            if (sx.side1 >= 0) {
                const DomainDev &n = ctx->doms[sx.side1];
                const long long rr = (long long)(q.row + sx.side1_offset) * n.nx + n.nx - sx.wlen;
                L.s1_px = n.base + n.f_px + rr;
                L.s1_py = n.base + n.f_py + rr;
            }
            if (sx.side2 >= 0) {
                const DomainDev &n = ctx->doms[sx.side2];
                const long long rr = (long long)(q.row + sx.side2_offset) * n.nx;
                L.s2_px = n.base + n.f_px + rr;
                L.s2_py = n.base + n.f_py + rr;
            }
            L.n_main = d.nx;
            L.w = sx.wlen;
            L.coef = it->second;
            L.dst = q.dst;
            lines.push_back(L);
        }
        SIPoint P{};
        try {
            P.coef = add_weights(sy, (double)rc.offset_y, cy);
        } catch (const std::exception &e) {
            ctx->err = std::string("spectral receiver: ") + e.what();
            return PSTD_ERR_INVALID;
        }
        P.K = c.ny + 2 * w;
        P.col = col0;
        P.rec = r;
        points.push_back(P);
        col_total += P.K;
    }
    if (points.empty()) return PSTD_OK;
    std::stable_sort(lines.begin(), lines.end(), [](const SILine &x, const SILine &y) { return x.m_px < y.m_px; });
    HIP_TRY(hipMalloc(&ctx->d_si_lines, lines.size() * sizeof(SILine)));
    HIP_TRY(hipMalloc(&ctx->d_si_points, points.size() * sizeof(SIPoint)));
    HIP_TRY(hipMalloc(&ctx->d_si_coef, coef.size() * sizeof(float)));
    HIP_TRY(hipMalloc(&ctx->d_si_col, (size_t)col_total * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(ctx->d_si_lines, lines.data(), lines.size() * sizeof(SILine), hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->d_si_points, points.data(), points.size() * sizeof(SIPoint), hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->d_si_coef, coef.data(), coef.size() * sizeof(float), hipMemcpyHostToDevice,
                           ctx->stream));
    // window rows of missing neighbours are never written: zeros
    HIP_TRY(hipMemsetAsync(ctx->d_si_col, 0, (size_t)col_total * sizeof(float), ctx->stream));
    ctx->n_si_lines = (int)lines.size();
    ctx->n_si = (int)points.size();
    return PSTD_OK;
}

static int upload_impl(pstd_ctx *ctx, const pstd_plan_desc *plan, const std::vector<int> *dom_scene = nullptr) {
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->release();
    ctx->doms.clear();
    ctx->tiles_small.clear();
    ctx->small_fmax = 0;
    ctx->small_nt = 256;
    ctx->fft_segs.clear();
    ctx->tiles_extra.clear();
    ctx->tiles_4096.clear();
    ctx->tiles_8192.clear();
    ctx->tiles_gemm.clear();
    ctx->max_ranges.clear();
    ctx->n_scenes = 1;
    if (dom_scene)
        for (int sc : *dom_scene) ctx->n_scenes = std::max(ctx->n_scenes, sc + 1);
    auto scene_of = [&](int d) { return dom_scene ? (*dom_scene)[d] : 0; };
    ctx->gemm_flops = 0;
    ctx->gemm_cells = 0;
    ctx->plan_domains.assign(plan->domains, plan->domains + plan->n_domains);
    // arena layout: one block per domain (vx, vy, px, py), 256 B aligned
    long long off = 0;
    auto align = [](long long v) { return (v + 63) & ~63LL; };
    for (int i = 0; i < plan->n_domains; ++i) {
        const pstd_plan_domain &p = plan->domains[i];
        if (p.nx <= 0 || p.ny <= 0) {
            ctx->err = "domain with empty extent";
            return PSTD_ERR_INVALID;
        }
        const long long cells = (long long)p.nx * p.ny, block = 4 * cells + p.nx + p.ny + 4 * 64;
        if (block > kMaxDomainFloats) {
            ctx->err = "domain of " + std::to_string(cells) + " cells exceeds one buffer block";
            return PSTD_ERR_UNSUPPORTED;
        }
        DomainDev d{};
        d.nx = p.nx;
        d.ny = p.ny;
        d.rho = p.rho;
        d.is_pml = p.is_pml;
        d.base = off;
        long long f = 0;
        d.f_vx = (int)f;
        f = align(f + (long long)p.ny * (p.nx + 1));
        d.f_vy = (int)f;
        f = align(f + (long long)(p.ny + 1) * p.nx);
        d.f_px = (int)f;
        f = align(f + cells);
        d.f_py = (int)f;
        f = align(f + cells + 1);  // a spare float after py (8-byte column accesses)
        d.size = (int)f;
        off += f;
        ctx->doms.push_back(d);
    }
    ctx->arena_floats = (size_t)std::max(off, 64LL);
    ctx->dt = plan->dt;
    ctx->c1 = plan->c1;
    for (int k = 0; k < 6; ++k) ctx->alpha[k] = plan->rk_alpha[k];
    // device plan cache: filters, twiddles, windows
    std::vector<float2> ctab;
    std::vector<float> ftab;
    std::map<std::tuple<int, int, int>, int> filt_cache;
    std::map<int, int> tw_cache;
    std::map<int, int> win_cache;
    const double dx = plan->dx;
    // source addressing of a domain's derivative input for (direction, field):
    // pressure reads p0 = px + py (row stride nx), velocity vx (X, stride nx+1) or vy (Y)
    // (field offsets within the domain block; velocity lines have no second array)
    auto src_addr = [&](const DomainDev &d, int dir, int field, int &a, int &b, int &ls, int &ps) {
        if (field == 0) {
            a = d.f_px;
            b = d.f_py;
            ls = dir == 0 ? d.nx : 1;
            ps = dir == 0 ? 1 : d.nx;
        } else if (dir == 0) {
            a = b = d.f_vx;
            ls = d.nx + 1;
            ps = 1;
        } else {
            a = b = d.f_vy;
            ls = 1;
            ps = d.nx;
        }
    };
    std::vector<std::vector<char>> covered(plan->n_domains, std::vector<char>(4, 0));
    std::vector<uint16_t> dmat;
    std::map<std::string, std::tuple<long long, int, int>> dcache;  // matrix key -> (hi offset, d_exp, dot row)
    for (int si = 0; si < plan->n_segments; ++si) {
        const pstd_plan_segment &s = plan->segments[si];
        if (s.domain < 0 || s.domain >= plan->n_domains || s.side1 >= plan->n_domains || s.side2 >= plan->n_domains ||
            (s.direction != 0 && s.direction != 1) || (s.field != 0 && s.field != 1) || s.wlen <= 0) {
            ctx->err = "invalid segment " + std::to_string(si);
            return PSTD_ERR_INVALID;
        }
        if (s.wlen > 32) {
            ctx->err = "window half-length " + std::to_string(s.wlen) + " above 32 (patch error above ~72 dB)";
            return PSTD_ERR_UNSUPPORTED;
        }
        const DomainDev &dm = ctx->doms[s.domain];
        const int n = s.direction == 0 ? dm.nx : dm.ny;
        const int nlines = s.direction == 0 ? dm.ny : dm.nx;
        if (s.line_begin < 0 || s.line_end > nlines || s.line_begin >= s.line_end) {
            ctx->err = "segment line range outside its domain";
            return PSTD_ERR_INVALID;
        }
        TileDesc g{};
        g.kind = 0;
        g.dir = s.direction;
        g.vel = s.field;
        g.wlen = s.wlen;
        g.n_main = s.field == 0 ? n : n + 1;
        g.n_out = s.field == 0 ? n + 1 : n;
        g.F = next_pow2(g.n_main + 2 * s.wlen);
        g.c_s1 = s.c_side1;
        g.c_m1 = s.c_main1;
        g.c_s2 = s.c_side2;
        g.c_m2 = s.c_main2;
        // update factor of Solver::update_field_values (kernel/Solver.cpp:174-181),
        // in double: 1/rho for the velocity fields, rho*c1^2 for the pressure fields
        g.coef = s.field == 0 ? (float)(1.0 / (double)dm.rho) : (float)((double)dm.rho * plan->c1 * plan->c1);
        g.is_pml = dm.is_pml;
        g.wlog = 0;
        while ((1 << g.wlog) < s.wlen) ++g.wlog;
        g.m_dom = g.s1_dom = g.s2_dom = dm.base;
        g.m_size = g.s1_size = g.s2_size = dm.size;
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
        src_addr(dm, s.direction, s.field, g.m_a, g.m_b, g.m_ls, g.m_ps);
        g.s1_len = g.s2_len = -1;
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
            if (side == 0) {
                src_addr(dn, s.direction, s.field, g.s1_a, g.s1_b, g.s1_ls, g.s1_ps);
                g.s1_line = o;
                g.s1_len = nlen;
                g.s1_dom = dn.base;
                g.s1_size = dn.size;
            } else {
                src_addr(dn, s.direction, s.field, g.s2_a, g.s2_b, g.s2_ls, g.s2_ps);
                g.s2_line = o;
                g.s2_len = nlen;
                g.s2_dom = dn.base;
                g.s2_size = dn.size;
            }
        }
        if (s.interp_only) {
            if (s.direction != 0 || s.field != 0) {
                ctx->err = "interpolation-only segment " + std::to_string(si) + " is not an X pressure segment";
                return PSTD_ERR_INVALID;
            }
            continue;  // read by the spectral receivers only
        }
        // output field: X pressure -> vx, X velocity -> px, Y pressure -> vy, Y velocity -> py
        if (s.direction == 0) {
            g.o_f = s.field == 0 ? dm.f_vx : dm.f_px;
            g.o_ls = s.field == 0 ? dm.nx + 1 : dm.nx;
            g.o_ps = 1;
        } else {
            g.o_f = s.field == 0 ? dm.f_vy : dm.f_py;
            g.o_ls = 1;
            g.o_ps = dm.nx;
        }
        covered[s.domain][s.direction == 0 ? (s.field == 0 ? 2 : 0) : (s.field == 0 ? 3 : 1)] = 1;
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
        g.tw2_off = 0;
        if (g.F == 8192) {
            // the split transform's two 4096-point halves
            auto t2 = tw_cache.find(4096);
            if (t2 == tw_cache.end()) {
                std::vector<float2> tw;
                build_twiddles(4096, tw);
                t2 = tw_cache.emplace(4096, (int)ctab.size()).first;
                ctab.insert(ctab.end(), tw.begin(), tw.end());
            }
            g.tw2_off = t2->second;
        }
        auto wt = win_cache.find(s.wlen);
        if (wt == win_cache.end()) {
            std::vector<float> w;
            build_window(s.wlen, (int)plan->patch_error, w);
            wt = win_cache.emplace(s.wlen, (int)ftab.size()).first;
            ftab.insert(ftab.end(), w.begin(), w.end());
        }
        g.win_off = wt->second;
        g.scene = scene_of(s.domain);
        // GEMM form: segments that fill the 128 x 256 tiles at least to 3/4
        // become GEMM tiles; thin ones (PML strips) stay FFT tiles
        const int nlines_seg = s.line_end - s.line_begin;
        if ((ctx->form == PSTD_FORM_GEMM_F16 || ctx->form == PSTD_FORM_GEMM_BF16) &&
            gemm_fills(nlines_seg, s.field == 0 ? g.n_out - 1 : g.n_out)) {
            int rc = add_gemm_tiles(ctx, plan, s, g, dmat, ftab, dcache);
            if (rc) return rc;
            continue;
        }
        g.line0 = s.line_begin;
        g.count = s.line_end - s.line_begin;
        if (g.F <= 2048) ctx->small_fmax = std::max(ctx->small_fmax, g.F);
        ctx->fft_segs.push_back(g);
    }
    // passthrough tiles for fields no segment updates (rigid domains, directions
    // openPSTD does not update).  covered[] index: 0 px, 1 py, 2 vx, 3 vy
    for (int i = 0; i < plan->n_domains; ++i) {
        const DomainDev &d = ctx->doms[i];
        const int offs[4] = {d.f_px, d.f_py, d.f_vx, d.f_vy};
        const int sizes[4] = {d.nx * d.ny, d.nx * d.ny, d.ny * (d.nx + 1), (d.ny + 1) * d.nx};
        for (int f = 0; f < 4; ++f) {
            if (covered[i][f]) continue;
            for (int e = 0; e < sizes[f]; e += kPassChunk) {
                TileDesc t{};
                t.kind = 1;
                t.line0 = e;
                t.count = std::min(kPassChunk, sizes[f] - e);
                t.m_dom = d.base;
                t.m_size = d.size;
                t.o_f = offs[f];
                t.is_pml = d.is_pml;
                t.scene = scene_of(i);
                t.vel = f < 2 ? 1 : 0;  // maxima class: px/py as written by velocity lines
                ctx->tiles_extra.push_back(t);
            }
        }
    }
    if (plan->n_receivers > 0) {
        TileDesc t{};
        t.kind = 2;  // receiver sampling of the previous step (first stage only)
        ctx->tiles_extra.push_back(t);
    }
    if (ctab.empty()) ctab.push_back(make_float2(0, 0));
    if (ftab.empty()) ftab.push_back(0);
    // PML factor arena (1 where no PML)
    std::vector<float> pml(ctx->arena_floats, 1.f);
    for (int i = 0; i < plan->n_domains; ++i) {
        const pstd_plan_domain &p = plan->domains[i];
        if (!p.is_pml) continue;
        const DomainDev &d = ctx->doms[i];
        const float *src[4] = {p.pml_vx, p.pml_vy, p.pml_px, p.pml_py};
        const long long offs[4] = {d.base + d.f_vx, d.base + d.f_vy, d.base + d.f_px, d.base + d.f_py};
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
    std::vector<ReceiverCell> cells;
    for (int r = 0; r < plan->n_receivers; ++r) {
        const pstd_plan_receiver &rc = plan->receivers[r];
        if (rc.domain < 0 || rc.domain >= plan->n_domains || rc.index < 0 ||
            rc.index >= plan->domains[rc.domain].nx * plan->domains[rc.domain].ny) {
            ctx->err = "receiver outside its domain";
            return PSTD_ERR_INVALID;
        }
        const DomainDev &d = ctx->doms[rc.domain];
        if (rc.spectral) cells.push_back({-1, -1});
        else cells.push_back({d.base + d.f_px + rc.index, d.base + d.f_py + rc.index});
    }
    ctx->nrec = (int)cells.size();
    {
        int rc = build_spectral_receivers(ctx, plan);
        if (rc) return rc;
    }
    std::vector<P0Range> ranges;
    for (const DomainDev &d : ctx->doms) ranges.push_back({d.base + d.f_px, d.base + d.f_py, (long long)d.nx * d.ny});
    ctx->n_ranges = (int)ranges.size();
    if (!ctx->tiles_gemm.empty()) {
        // what a first stage after host writes reads: p0 (P0I) and vx, vy (TN)
        for (int i = 0; i < plan->n_domains; ++i) {
            const DomainDev &d = ctx->doms[i];
            const int sc = scene_of(i);
            ctx->max_ranges.push_back({3LL * (long long)ctx->arena_floats + d.base + d.f_px, (long long)d.nx * d.ny, 2 * sc});
            ctx->max_ranges.push_back({d.base + d.f_vx, (long long)d.ny * (d.nx + 1), 2 * sc + 1});
            ctx->max_ranges.push_back({d.base + d.f_vy, (long long)(d.ny + 1) * d.nx, 2 * sc + 1});
        }
    }
    // allocate + upload
    const size_t af = ctx->arena_floats;
    HIP_TRY(hipMalloc(&ctx->arena, 4 * af * sizeof(float)));
    HIP_TRY(hipMemsetAsync(ctx->arena, 0, 4 * af * sizeof(float), ctx->stream));
    HIP_TRY(hipMalloc(&ctx->pml, af * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(ctx->pml, pml.data(), af * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMalloc(&ctx->ctab, ctab.size() * sizeof(float2)));
    HIP_TRY(hipMemcpyAsync(ctx->ctab, ctab.data(), ctab.size() * sizeof(float2), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMalloc(&ctx->ftab, ftab.size() * sizeof(float)));
    HIP_TRY(hipMemcpyAsync(ctx->ftab, ftab.data(), ftab.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    {
        const int rc = upload_tiles(ctx);
        if (rc) return rc;
    }
    if (!ctx->tiles_gemm.empty()) {
        // segment order (a scene's segments together, a line block's output
        // tiles adjacent): measured faster than grouping the tiles by matrix
        HIP_TRY(hipMalloc(&ctx->d_gemm, ctx->tiles_gemm.size() * sizeof(GemmTile)));
        HIP_TRY(hipMemcpyAsync(ctx->d_gemm, ctx->tiles_gemm.data(), ctx->tiles_gemm.size() * sizeof(GemmTile),
                               hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMalloc(&ctx->dmat, std::max<size_t>(1, dmat.size()) * sizeof(uint16_t)));
        HIP_TRY(hipMemcpyAsync(ctx->dmat, dmat.data(), dmat.size() * sizeof(uint16_t), hipMemcpyHostToDevice,
                               ctx->stream));
        const size_t nmax = 3 * 2 * (size_t)ctx->n_scenes;
        HIP_TRY(hipMalloc(&ctx->umax, nmax * sizeof(unsigned)));
        HIP_TRY(hipMemsetAsync(ctx->umax, 0, nmax * sizeof(unsigned), ctx->stream));
        HIP_TRY(hipMalloc(&ctx->d_max_ranges, ctx->max_ranges.size() * sizeof(MaxRange)));
        HIP_TRY(hipMemcpyAsync(ctx->d_max_ranges, ctx->max_ranges.data(), ctx->max_ranges.size() * sizeof(MaxRange),
                               hipMemcpyHostToDevice, ctx->stream));
    }
    HIP_TRY(hipMalloc(&ctx->d_cells, std::max<size_t>(1, cells.size()) * sizeof(ReceiverCell)));
    if (!cells.empty())
        HIP_TRY(hipMemcpyAsync(ctx->d_cells, cells.data(), cells.size() * sizeof(ReceiverCell), hipMemcpyHostToDevice,
                               ctx->stream));
    HIP_TRY(hipMalloc(&ctx->d_ranges, std::max<size_t>(1, ranges.size()) * sizeof(P0Range)));
    if (!ranges.empty())
        HIP_TRY(hipMemcpyAsync(ctx->d_ranges, ranges.data(), ranges.size() * sizeof(P0Range), hipMemcpyHostToDevice,
                               ctx->stream));
    ctx->sample_cap = 4096;
    HIP_TRY(hipMalloc(&ctx->d_samples, (size_t)ctx->sample_cap * std::max(1, ctx->nrec) * sizeof(float)));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->steps_done = 0;
    ctx->p0_explicit = true;
    ctx->planned = true;
    const int rc = autotune(ctx);
    if (rc) ctx->planned = false;  // a failed trial run leaves no usable launch state
    return rc;
}

// PSTD_FORM_AUTO: the plan uploaded (and autotuned) in the FFT form and in the
// FP16 GEMM form, the one with the shorter tuned stage kept; the choice is
// remembered per (device, plan signature), so a later upload of a plan of
// the same geometry uploads once.
static int upload_form(pstd_ctx *ctx, const pstd_plan_desc *plan, const std::vector<int> *dom_scene) {
    if (ctx->form_req != PSTD_FORM_AUTO) {
        ctx->form = ctx->form_req;
        return upload_impl(ctx, plan, dom_scene);
    }
    ctx->form = PSTD_FORM_FFT;
    int rc = upload_impl(ctx, plan, dom_scene);
    if (rc) return rc;
    const std::pair<int, uint64_t> key(ctx->device, plan_signature(ctx));
    int chosen = -1;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_form.find(key);
        if (it != g_form.end()) chosen = it->second;
    }
    if (chosen < 0) {
        const float fft_ms = ctx->tuned_stage_ms;
        const int trials = ctx->trial_steps;
        ctx->form = PSTD_FORM_GEMM_F16;
        rc = upload_impl(ctx, plan, dom_scene);
        if (rc) return rc;
        const bool gemm_better = !ctx->tiles_gemm.empty() && ctx->tuned_stage_ms < fft_ms;
        chosen = gemm_better ? PSTD_FORM_GEMM_F16 : PSTD_FORM_FFT;
        ctx->trial_steps += trials;
        {
            std::lock_guard<std::mutex> lock(g_tune_mu);
            g_form[key] = chosen;
        }
        if (gemm_better) return PSTD_OK;
        const int t = ctx->trial_steps;
        ctx->form = PSTD_FORM_FFT;
        rc = upload_impl(ctx, plan, dom_scene);  // its launch shape comes from the autotune cache
        ctx->trial_steps += t;
        return rc;
    }
    if (chosen == PSTD_FORM_FFT) return PSTD_OK;
    ctx->form = chosen;
    return upload_impl(ctx, plan, dom_scene);
}

int pstd_ctx_upload_plan(pstd_ctx *ctx, const pstd_plan_desc *plan) {
    if (!ctx || !plan) return PSTD_ERR_INVALID;
    return upload_form(ctx, plan, nullptr);
}

int pstd_ctx_upload_plans(pstd_ctx *ctx, const pstd_plan_desc *plans, int n_plans) {
    if (!ctx || !plans || n_plans <= 0) return PSTD_ERR_INVALID;
    if (n_plans == 1) return upload_form(ctx, plans, nullptr);
    // independent scenes: one plan with the domain lists concatenated; scenes
    // share the time step, the RK coefficients and the grid (one launch per
    // stage advances all of them)
    std::vector<pstd_plan_domain> doms;
    std::vector<pstd_plan_segment> segs;
    std::vector<pstd_plan_receiver> rcv;
    std::vector<int> dom_scene;
    const pstd_plan_desc &p0 = plans[0];
    for (int i = 0; i < n_plans; ++i) {
        const pstd_plan_desc &p = plans[i];
        bool same = p.dx == p0.dx && p.dt == p0.dt && p.c1 == p0.c1 && p.patch_error == p0.patch_error;
        for (int k = 0; k < 6; ++k) same = same && p.rk_alpha[k] == p0.rk_alpha[k];
        if (!same) {
            ctx->err = "batched scenes must share dx, dt, c1, patch error and RK coefficients (scene " +
                       std::to_string(i) + " differs)";
            return PSTD_ERR_INVALID;
        }
        if (p.n_domains < 0 || p.n_segments < 0 || p.n_receivers < 0) return PSTD_ERR_INVALID;
        const int base = (int)doms.size();
        for (int d = 0; d < p.n_domains; ++d) {
            doms.push_back(p.domains[d]);
            dom_scene.push_back(i);
        }
        for (int k = 0; k < p.n_segments; ++k) {
            pstd_plan_segment s = p.segments[k];
            const bool bad = s.domain < 0 || s.domain >= p.n_domains || s.side1 >= p.n_domains || s.side2 >= p.n_domains;
            if (bad) {
                ctx->err = "segment of scene " + std::to_string(i) + " names a domain outside its scene";
                return PSTD_ERR_INVALID;
            }
            s.domain += base;
            if (s.side1 >= 0) s.side1 += base;
            if (s.side2 >= 0) s.side2 += base;
            segs.push_back(s);
        }
        for (int r = 0; r < p.n_receivers; ++r) {
            pstd_plan_receiver c = p.receivers[r];
            if (c.domain < 0 || c.domain >= p.n_domains) {
                ctx->err = "receiver of scene " + std::to_string(i) + " outside its scene";
                return PSTD_ERR_INVALID;
            }
            c.domain += base;
            rcv.push_back(c);
        }
    }
    pstd_plan_desc m = p0;
    m.n_domains = (int)doms.size();
    m.domains = doms.data();
    m.n_segments = (int)segs.size();
    m.segments = segs.data();
    m.n_receivers = (int)rcv.size();
    m.receivers = rcv.data();
    return upload_form(ctx, &m, &dom_scene);
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
    if (ctx->umax && field != PSTD_FIELD_P0) {
        // GEMM form: the next first stage recomputes the scaling maxima from
        // the state, which needs p0 in its explicit array
        int rc = materialise_p0(ctx);
        if (rc) return rc;
    }
    if (field == PSTD_FIELD_P0) {
        int rc = materialise_p0(ctx);
        if (rc) return rc;
        dst = ctx->P0I() + field_off(ctx->doms[domain], PSTD_FIELD_PX);
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
            HIP_TRY(hipMemcpyAsync(host, ctx->P0I() + field_off(d, PSTD_FIELD_PX), count * sizeof(float), hipMemcpyDeviceToHost,
                                   ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
        } else {
            std::vector<float> py(count);
            HIP_TRY(hipMemcpyAsync(host, ctx->S(1) + field_off(d, PSTD_FIELD_PX), count * sizeof(float), hipMemcpyDeviceToHost,
                                   ctx->stream));
            HIP_TRY(hipMemcpyAsync(py.data(), ctx->S(1) + field_off(d, PSTD_FIELD_PY), count * sizeof(float), hipMemcpyDeviceToHost,
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

// Captures enqueue_steps(graph_steps) into a graph.  Only valid once p0 comes
// from the stage outputs (after the first step), which every replay keeps.
static int build_graph(pstd_ctx *ctx) {
    ctx->drop_graph();
    if (ctx->nrec > 0)
        HIP_TRY(hipMalloc(&ctx->d_gsamples, (size_t)ctx->graph_steps * ctx->nrec * sizeof(float)));
    float *const samples = ctx->d_samples;
    ctx->d_samples = ctx->d_gsamples;  // captured: the graph's own sample slots 0 .. graph_steps-1
    HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_steps(ctx, ctx->graph_steps);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    ctx->d_samples = samples;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        ctx->err = std::string("graph capture: ") + hipGetErrorString(e);
        return PSTD_ERR_DEVICE;
    }
    ctx->graph = g;
    HIP_TRY(hipGraphInstantiate(&ctx->graph_exec, ctx->graph, nullptr, nullptr, 0));
    return PSTD_OK;
}

static int advance_impl(pstd_ctx *ctx, int steps, float *samples) {
    if (!ctx || !ctx->planned) return PSTD_ERR_INVALID;
    if (steps < 0) return PSTD_ERR_INVALID;
    HIP_TRY(hipSetDevice(ctx->device));
    if (ctx->max_stale) {
        // the next first stage recomputes the maxima from the state (explicit p0)
        const int rc = materialise_p0(ctx);
        if (rc) return rc;
        ctx->max_stale = false;
    }
    int done = 0;
    const int gs = ctx->graph_steps;
    while (done < steps) {
        // one chunk of at most sample_cap steps fills sample slots 0 .. chunk-1:
        // graph replays where a whole graph fits (each copies its samples into
        // place on the device), eager launches for the rest and for a first step
        // after host writes (the graph reads p0 from the stage outputs)
        const int chunk = std::min(steps - done, ctx->sample_cap);
        int pos = 0;
        while (pos < chunk) {
            const int left = chunk - pos;
            if (gs > 0 && !ctx->p0_explicit && left >= gs) {
                if (!ctx->graph_exec) {
                    int rc = build_graph(ctx);
                    if (rc) return rc;
                }
                HIP_TRY(hipGraphLaunch(ctx->graph_exec, ctx->stream));
                if (samples && ctx->nrec > 0)
                    HIP_TRY(hipMemcpyAsync(ctx->d_samples + (size_t)pos * ctx->nrec, ctx->d_gsamples,
                                           (size_t)gs * ctx->nrec * sizeof(float), hipMemcpyDeviceToDevice,
                                           ctx->stream));
                pos += gs;
            } else {
                const int n = (gs > 0 && ctx->p0_explicit && left > gs) ? 1 : (gs > 0 && left >= gs ? gs : left);
                int rc = enqueue_steps(ctx, n, pos);
                if (rc) return rc;
                pos += n;
            }
        }
        ctx->steps_done += chunk;
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

int pstd_ctx_set_stage_overlap(pstd_ctx *ctx, int on) {
    if (!ctx) return PSTD_ERR_INVALID;
    if (on && !(ctx->side && ctx->ev_fork && ctx->ev_join)) {
        ctx->err = "no side stream for overlapped stages";
        return PSTD_ERR_DEVICE;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->drop_graph();
    ctx->overlap = on != 0;
    return PSTD_OK;
}

int pstd_ctx_capture_graph(pstd_ctx *ctx) {
    if (!ctx || !ctx->planned) return PSTD_ERR_INVALID;
    if (ctx->graph_steps <= 0 || ctx->graph_exec) return PSTD_OK;
    if (ctx->p0_explicit) {
        ctx->err = "the step graph is captured once p0 comes from the stage outputs (after a time step)";
        return PSTD_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return build_graph(ctx);
}

int pstd_ctx_set_graph_steps(pstd_ctx *ctx, int steps) {
    if (!ctx || steps < 0) return PSTD_ERR_INVALID;
    if (ctx->planned && steps > ctx->sample_cap) {
        ctx->err = "graph longer than the sample buffer (" + std::to_string(ctx->sample_cap) + " steps)";
        return PSTD_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->drop_graph();
    ctx->graph_steps = steps;
    return PSTD_OK;
}

int pstd_ctx_set_form(pstd_ctx *ctx, int form) {
    if (!ctx || (form != PSTD_FORM_FFT && form != PSTD_FORM_GEMM_F16 && form != PSTD_FORM_GEMM_BF16 &&
                 form != PSTD_FORM_AUTO))
        return PSTD_ERR_INVALID;
    if (ctx->planned) {
        ctx->err = "the derivative form is chosen before the plan is uploaded";
        return PSTD_ERR_INVALID;
    }
    ctx->form_req = form;
    ctx->form = form == PSTD_FORM_AUTO ? PSTD_FORM_FFT : form;
    return PSTD_OK;
}

int pstd_ctx_get_form(const pstd_ctx *ctx) { return ctx ? ctx->form : PSTD_ERR_INVALID; }

int pstd_ctx_set_kernel_flavour(pstd_ctx *ctx, int flavour) {
    if (!ctx || flavour < kEarly || flavour > kMid) return PSTD_ERR_INVALID;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->drop_graph();
    ctx->lean = flavour;
    return PSTD_OK;
}

int pstd_ctx_get_tuning(const pstd_ctx *ctx, int32_t *lean, int32_t *xcd_contiguous, double *stage_ms) {
    if (!ctx || !ctx->planned) return PSTD_ERR_INVALID;
    if (lean) *lean = ctx->lean;
    if (xcd_contiguous) *xcd_contiguous = ctx->xcd_remap;
    if (stage_ms) *stage_ms = ctx->tuned_stage_ms;
    return PSTD_OK;
}

int pstd_ctx_get_autotune(const pstd_ctx *ctx, int32_t *trial_steps, int32_t *from_cache) {
    if (!ctx || !ctx->planned) return PSTD_ERR_INVALID;
    if (trial_steps) *trial_steps = ctx->trial_steps;
    if (from_cache) *from_cache = ctx->tune_cached ? 1 : 0;
    return PSTD_OK;
}

int pstd_ctx_set_tile_threads(pstd_ctx *ctx, int threads) {
    if (!ctx || !ctx->planned || (threads != 128 && threads != 256 && threads != 512)) return PSTD_ERR_INVALID;
    if (threads == 128 && ctx->small_fmax > 1024) {
        ctx->err = "128-thread tiles need every line of the F <= 2048 class to have F <= 1024";
        return PSTD_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->drop_graph();
    ctx->small_nt = threads;
    const int rc = upload_tiles(ctx);
    if (rc) {
        ctx->planned = false;  // the tile lists are gone with the failed upload
        return rc;
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PSTD_OK;
}

int pstd_ctx_get_tile_threads(const pstd_ctx *ctx) {
    if (!ctx || !ctx->planned) return PSTD_ERR_INVALID;
    return ctx->small_nt;
}

int pstd_ctx_set_tile_order(pstd_ctx *ctx, int xcd_contiguous) {
    if (!ctx) return PSTD_ERR_INVALID;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->drop_graph();
    ctx->xcd_remap = xcd_contiguous ? 1 : 0;
    return PSTD_OK;
}

int pstd_ctx_profile_stages(pstd_ctx *ctx, int steps, double *mean_stage_ms, int *launches) {
    if (!ctx || !ctx->planned || steps <= 0 || !mean_stage_ms) return PSTD_ERR_INVALID;
    HIP_TRY(hipSetDevice(ctx->device));
    hipEvent_t ev[2];
    HIP_TRY(hipEventCreate(&ev[0]));
    HIP_TRY(hipEventCreate(&ev[1]));
    int rc = PSTD_OK;
    float ms = 0;
    if (hipEventRecord(ev[0], ctx->stream) != hipSuccess) rc = PSTD_ERR_DEVICE;
    if (rc == PSTD_OK) rc = advance_impl(ctx, steps, nullptr);
    if (rc == PSTD_OK && hipEventRecord(ev[1], ctx->stream) != hipSuccess) rc = PSTD_ERR_DEVICE;
    if (rc == PSTD_OK && hipEventSynchronize(ev[1]) != hipSuccess) rc = PSTD_ERR_DEVICE;
    if (rc == PSTD_OK && hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) rc = PSTD_ERR_DEVICE;
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc) return rc;
    *mean_stage_ms = (double)ms / (6.0 * steps);
    if (launches) *launches = 6 * steps;
    return PSTD_OK;
}

int pstd_ctx_advance_static(pstd_ctx *ctx, int steps) {
    if (!ctx || !ctx->planned || steps < 0) return PSTD_ERR_INVALID;
    const bool gemm = !ctx->tiles_gemm.empty();
    if (gemm && ctx->n_scenes > kStaticScenes) {
        ctx->err = "allocation-free advance: " + std::to_string(ctx->n_scenes) + " scenes, at most " +
                   std::to_string(kStaticScenes);
        return PSTD_ERR_INVALID;
    }
    HIP_TRY(hipSetDevice(ctx->device));
    unsigned *sp = nullptr;
    if (gemm) {
        if (resolve_scratch(ctx->device, &sp) != PSTD_OK) {
            ctx->err = "allocation-free advance: no device address for the scratch array";
            return PSTD_ERR_DEVICE;
        }
        // the shared array starts from zero and the maxima of the current state
        // (explicit p0, then the first stage's reduction), as after host writes
        int rc = materialise_p0(ctx);
        if (rc) return rc;
        const int n = 3 * 2 * ctx->n_scenes;
        hipLaunchKernelGGL(fill_zero_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, sp, n);
        HIP_TRY(hipGetLastError());
    }
    unsigned *own = ctx->umax;
    if (gemm) ctx->umax = sp;
    int rc = PSTD_OK, done = 0;
    while (rc == PSTD_OK && done < steps) {  // eager launches only: nothing to instantiate
        const int chunk = std::min(steps - done, ctx->sample_cap);
        rc = enqueue_steps(ctx, chunk);
        if (rc == PSTD_OK) {
            ctx->steps_done += chunk;
            done += chunk;
        }
    }
    ctx->umax = own;
    if (gemm) ctx->max_stale = true;
    return rc;
}

int pstd_ctx_profile_kernels(pstd_ctx *ctx, int steps, double *ms, int *launches) {
    if (!ctx || !ctx->planned || steps <= 0 || !ms || !launches) return PSTD_ERR_INVALID;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    hipEvent_t ev[2];
    HIP_TRY(hipEventCreate(&ev[0]));
    HIP_TRY(hipEventCreate(&ev[1]));
    const int graph_steps = ctx->graph_steps;
    ctx->graph_steps = 0;  // eager launches: events cannot be recorded into the replayed graph
    ctx->prof = true;
    ctx->prof_events.clear();
    int rc = PSTD_OK;
    float total = 0;
    if (hipEventRecord(ev[0], ctx->stream) != hipSuccess) rc = PSTD_ERR_DEVICE;
    if (rc == PSTD_OK) rc = advance_impl(ctx, steps, nullptr);
    if (rc == PSTD_OK && hipEventRecord(ev[1], ctx->stream) != hipSuccess) rc = PSTD_ERR_DEVICE;
    if (rc == PSTD_OK && hipEventSynchronize(ev[1]) != hipSuccess) rc = PSTD_ERR_DEVICE;
    if (rc == PSTD_OK && hipEventElapsedTime(&total, ev[0], ev[1]) != hipSuccess) rc = PSTD_ERR_DEVICE;
    ctx->prof = false;
    ctx->graph_steps = graph_steps;
    for (int c = 0; c <= kClasses; ++c) {
        ms[c] = 0;
        launches[c] = 0;
    }
    for (auto &pe : ctx->prof_events) {
        float t = 0;
        if (rc == PSTD_OK && hipEventElapsedTime(&t, pe.second.first, pe.second.second) != hipSuccess) rc = PSTD_ERR_DEVICE;
        ms[pe.first] += t;
        launches[pe.first] += 1;
        (void)hipEventDestroy(pe.second.first);
        (void)hipEventDestroy(pe.second.second);
    }
    ctx->prof_events.clear();
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (rc) {
        if (ctx->err.empty()) ctx->err = "profile_kernels: HIP event failure";
        return rc;
    }
    ms[kClasses] = total;
    launches[kClasses] = 6 * steps;
    return PSTD_OK;
}

int pstd_ctx_gemm_work(const pstd_ctx *ctx, double *flops_per_stage, long long *outputs_per_stage) {
    if (!ctx || !ctx->planned) return PSTD_ERR_INVALID;
    if (flops_per_stage) *flops_per_stage = ctx->gemm_flops;
    if (outputs_per_stage) *outputs_per_stage = ctx->gemm_cells;
    return PSTD_OK;
}

long long pstd_ctx_cells_per_stage(const pstd_ctx *ctx) {
    if (!ctx || !ctx->planned) return -1;
    long long n = 0;
    for (const pstd_plan_domain &d : ctx->plan_domains)
        if (!d.is_rigid) n += (long long)d.nx * d.ny;
    return n;
}

}  // extern "C"
