// Register/LDS Stockham FFT building blocks for the line-derivative kernels (gfx950).
//
// A "line pair" is two real lines packed as one complex sequence
// z = x1 + i*x2 of length F (F a power of two, 128..8192).  T = F/8 threads
// work on one pair; thread t holds the 8 elements t + m*T (m = 0..7) in
// registers ("natural layout") at the start and end of each transform.  A
// Stockham radix-R pass (R in {2, 4, 8}) does 8/R butterflies per thread and
// exchanges through LDS; the first and last pass of every transform are
// radix 8, so the forward transform ends and the inverse transform starts in
// natural layout without touching LDS.
#pragma once

#include <hip/hip_runtime.h>

namespace pstdamd {
namespace dev {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// Explicit FMAs (the device code is built without implicit contraction, so
// the rounding is the same in every instantiation).
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_i(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2 &a0, float2 &a1) {
    float2 t = a0;
    a0 = cadd(t, a1);
    a1 = csub(t, a1);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2 &a0, float2 &a1, float2 &a2, float2 &a3) {
    float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_i<INV>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft8(float2 *a) {
    // even/odd split: E = DFT4(a0,a2,a4,a6), O = DFT4(a1,a3,a5,a7)
    dft4<INV>(a[0], a[2], a[4], a[6]);
    dft4<INV>(a[1], a[3], a[5], a[7]);
    const float c = 0.70710678118654752440f;
    // W8^1 O1, W8^2 O2, W8^3 O3 with W8 = exp(-+2 pi i / 8)
    float2 o1 = a[3], o2 = a[5], o3 = a[7];
    float2 w1 = INV ? make_float2(c * (o1.x - o1.y), c * (o1.x + o1.y)) : make_float2(c * (o1.x + o1.y), c * (o1.y - o1.x));
    float2 w2 = mul_i<INV>(o2);
    float2 w3 = INV ? make_float2(-c * (o3.x + o3.y), c * (o3.x - o3.y)) : make_float2(c * (o3.y - o3.x), -c * (o3.x + o3.y));
    float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6], o0 = a[1];
    a[0] = cadd(e0, o0);
    a[4] = csub(e0, o0);
    a[1] = cadd(e1, w1);
    a[5] = csub(e1, w1);
    a[2] = cadd(e2, w2);
    a[6] = csub(e2, w2);
    a[3] = cadd(e3, w3);
    a[7] = csub(e3, w3);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2 *a) {
    if constexpr (R == 2) dft2<INV>(a[0], a[1]);
    else if constexpr (R == 4) dft4<INV>(a[0], a[1], a[2], a[3]);
    else dft8<INV>(a);
}

// Pass schedule: radix of pass p for length F (first and last are 8).
__host__ __device__ constexpr int fft_npasses(int F) {
    return F == 128 ? 3 : F == 256 ? 3 : F == 512 ? 3 : F == 1024 ? 4 : F == 2048 ? 4 : F == 4096 ? 4 : 5;
}
__host__ __device__ constexpr int fft_radix(int F, int p) {
    return (p == 0 || p == fft_npasses(F) - 1) ? 8
           : F == 128  ? 2
           : F == 256  ? 4
           : F == 512  ? 8
           : F == 1024 ? 4
           : F == 2048 ? (p == 1 ? 8 : 4)
           : F == 4096 ? 8
           : (p == 1 ? 2 : 8);
}
__host__ __device__ constexpr int fft_span(int F, int p) {  // product of radices before pass p
    return p == 0 ? 1 : fft_span(F, p - 1) * fft_radix(F, p - 1);
}

// LDS address of element idx of pair p.  LAYOUT 0: pair-major (lanes vary the
// element index); LAYOUT 1: element-major (lanes vary the pair index).  Both
// make the natural-layout loads of a wavefront contiguous.  One pad slot per 8
// elements: the radix-8 passes write elements 8t + r, a 64 B lane stride that
// would put every lane of a wavefront on the same few LDS banks.
__host__ __device__ constexpr int lds_padded(int n) { return n + (n >> 3); }
template <int F, int P, int LAYOUT>
struct LdsMap {
    static constexpr int SIZE = P * lds_padded(F);
    __device__ __forceinline__ static int at(int p, int idx) {
        const int e = idx + (idx >> 3);
        return LAYOUT == 0 ? p * lds_padded(F) + e : e * P + p;
    }
};

// Twiddle registers: pass p >= 1 needs (R_p - 1) factors per butterfly and
// 8/R_p butterflies per thread.  They depend only on t, so a thread loads all
// of them once, before the transforms (the inverse uses their conjugates).
__host__ __device__ constexpr int fft_tw_before(int F, int p) {
    return p <= 1 ? 0 : fft_tw_before(F, p - 1) + (fft_radix(F, p - 1) - 1) * (8 / fft_radix(F, p - 1));
}
__host__ __device__ constexpr int fft_tw_total(int F) { return fft_tw_before(F, fft_npasses(F)); }

template <int F, int PASS = 1>
__device__ __forceinline__ void fft_load_twiddles(float2 *w, int t, const float2 *__restrict__ tw) {
    if constexpr (PASS < fft_npasses(F)) {
        constexpr int T = F / 8;
        constexpr int R = fft_radix(F, PASS);
        constexpr int NS = fft_span(F, PASS);
        constexpr int NB = 8 / R;
        constexpr int base = fft_tw_before(F, PASS);
#pragma unroll
        for (int s = 0; s < NB; ++s) {
            const int k = (t + s * T) & (NS - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) w[base + s * (R - 1) + (r - 1)] = tw[(r * k) * (F / (NS * R))];
        }
        fft_load_twiddles<F, PASS + 1>(w, t, tw);
    }
}

// One Stockham pass.  v: natural layout input.  Non-last passes write their
// outputs to `out`; the last pass leaves them in v (natural layout).
// TWG = false: w holds the thread's preloaded twiddle registers
// (fft_load_twiddles, or with PERPASS only this pass's, fft_pass_twiddles);
// TWG = true: w is the twiddle table exp(-2 pi i k/F) in global memory and
// the pass reads the factors it needs when it needs them.
template <int F, int P, int LAYOUT, int PASS, bool INV, bool TWG, bool PERPASS = false>
__device__ __forceinline__ void fft_pass(float2 (&v)[8], int p, int t, float2 *out, const float2 *w) {
    constexpr int T = F / 8;
    constexpr int R = fft_radix(F, PASS);
    constexpr int NS = fft_span(F, PASS);
    constexpr int NB = 8 / R;
    constexpr bool LAST = PASS == fft_npasses(F) - 1;
    constexpr int base_tw = fft_tw_before(F, PASS);
    using L = LdsMap<F, P, LAYOUT>;
#pragma unroll
    for (int s = 0; s < NB; ++s) {
        const int j = t + s * T;
        const int k = j & (NS - 1);
        float2 a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = v[s + r * NB];
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 wv = TWG ? w[(r * k) * (F / (NS * R))]
                                : w[(PERPASS ? 0 : base_tw) + s * (R - 1) + (r - 1)];
                if (INV) wv.y = -wv.y;
                a[r] = cmul(a[r], wv);
            }
        }
        dft<R, INV>(a);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = a[r];
        } else {
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) out[L::at(p, base + r * NS)] = a[r];
        }
    }
}

template <int F, int P, int LAYOUT>
__device__ __forceinline__ void lds_load(float2 (&v)[8], int p, int t, const float2 *in) {
    constexpr int T = F / 8;
    using L = LdsMap<F, P, LAYOUT>;
#pragma unroll
    for (int m = 0; m < 8; ++m) v[m] = in[L::at(p, t + m * T)];
}

// Twiddles of one pass for thread t, read from the table exp(-2 pi i k/F):
// NB butterflies x (R-1) factors, at most 7.
constexpr int kTwPass = 7;
template <int F, int PASS>
__device__ __forceinline__ void fft_pass_twiddles(float2 (&w)[kTwPass], int t, const float2 *__restrict__ tw) {
    if constexpr (PASS >= 1 && PASS < fft_npasses(F)) {
        constexpr int T = F / 8;
        constexpr int R = fft_radix(F, PASS);
        constexpr int NS = fft_span(F, PASS);
        constexpr int NB = 8 / R;
#pragma unroll
        for (int s = 0; s < NB; ++s) {
            const int k = (t + s * T) & (NS - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) w[s * (R - 1) + (r - 1)] = tw[(r * k) * (F / (NS * R))];
        }
    }
}

// Exchange point after pass PASS.  Rows of F <= 512 (T = F/8 <= 64 threads
// per pair) keep every line pair inside one wavefront, and a pair's LDS slots
// (in both buffers, and its staged raw lines) are its own; the LDS executes
// one wavefront's accesses in issue order, so a compiler fence is enough
// there.  Everything else synchronises the workgroup.
template <int F, int LAYOUT, bool INV, int PASS>
__device__ __forceinline__ void pass_sync() {
    if constexpr (LAYOUT == 0 && F / 8 <= 64) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    else __syncthreads();
}

// Transform with the twiddles of each pass loaded one pass ahead, into
// registers: the loads of pass p+1 are in flight while pass p computes and
// exchanges.  wcur: this pass's twiddles (unused by pass 0).
template <int F, int P, int LAYOUT, bool INV, int PASS = 0>
__device__ __forceinline__ void fft_run_ahead(float2 (&v)[8], int p, int t, float2 *buf0, float2 *buf1,
                                              const float2 *tw, const float2 (&wcur)[kTwPass]) {
    constexpr int NP = fft_npasses(F);
    float2 wnext[kTwPass];
    if constexpr (PASS + 1 < NP) fft_pass_twiddles<F, PASS + 1>(wnext, t, tw);
    float2 *out = (PASS & 1) ? buf1 : buf0;
    fft_pass<F, P, LAYOUT, PASS, INV, false, true>(v, p, t, out, wcur);
    if constexpr (PASS + 1 < NP) {
        pass_sync<F, LAYOUT, INV, PASS>();
        lds_load<F, P, LAYOUT>(v, p, t, out);
        fft_run_ahead<F, P, LAYOUT, INV, PASS + 1>(v, p, t, buf0, buf1, tw, wnext);
    }
}

// Full transform (forward if !INV) on natural-layout registers, through the two
// LDS buffers.  Every thread of the block must call it (it synchronises).
template <int F, int P, int LAYOUT, bool INV, bool TWG, int PASS = 0>
__device__ __forceinline__ void fft_run(float2 (&v)[8], int p, int t, float2 *buf0, float2 *buf1, const float2 *w) {
    constexpr int NP = fft_npasses(F);
    float2 *out = (PASS & 1) ? buf1 : buf0;
    fft_pass<F, P, LAYOUT, PASS, INV, TWG>(v, p, t, out, w);
    if constexpr (PASS + 1 < NP) {
        pass_sync<F, LAYOUT, INV, PASS>();
        lds_load<F, P, LAYOUT>(v, p, t, out);
        fft_run<F, P, LAYOUT, INV, TWG, PASS + 1>(v, p, t, buf0, buf1, w);
    }
}


}  // namespace dev
}  // namespace pstdamd
