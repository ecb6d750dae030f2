// Device plan cache tables (see plan_tables.h).
#include "plan_tables.h"

#include <cmath>
#include <stdexcept>

namespace pstdamd {

std::vector<std::complex<double>> derivative_filter(int F, int L, int field, double dx) {
    std::vector<std::complex<double>> out(F, 0.0);
    const double dk = (M_PI / dx) / (L / 2);
    std::vector<std::complex<double>> d(F / 2 + 1);
    for (int m = 0; m <= F / 2; ++m) {
        const double k = m * dk;
        d[m] = std::exp(std::complex<double>(0, (field == 0 ? -1 : 1) * k * dx * 0.5)) * std::complex<double>(0, k);
    }
    d[0] = 0;
    d[F / 2] = d[F / 2].real();
    for (int m = 0; m <= F / 2; ++m) {
        const std::complex<double> h = d[m] / (double)F;
        out[m] = h;
        if (m > 0 && m < F / 2) out[F - m] = std::conj(h);
    }
    return out;
}

std::vector<std::complex<double>> shift_filter(int F, int L, double delta) {
    std::vector<std::complex<double>> out(F, 0.0);
    for (int m = 0; m <= F / 2; ++m) {
        std::complex<double> h = std::polar(1.0, M_PI * m * delta / (L / 2)) / (double)F;
        if (m == F / 2) h = h.real();
        out[m] = h;
        if (m > 0 && m < F / 2) out[F - m] = std::conj(h);
    }
    return out;
}

std::vector<double> window_coefficients(int w, double patch_error) {
    std::vector<double> out(2 * w + 1);
    const double alpha = (patch_error - 40) / 20.0 + 1;
    for (int i = 0; i <= 2 * w; ++i) {
        const double x = (double)(i - w) / w;
        out[i] = std::exp(-std::log(10.0) * alpha * std::pow(x, 6));
    }
    return out;
}

namespace {

// In-place iterative radix-2 DFT, sign +1: a[r] <- sum_k a[k] exp(+2 pi i k r / n).
void idft_pow2(std::vector<std::complex<double>> &a) {
    const int n = (int)a.size();
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const double ang = 2 * M_PI / len;
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < len / 2; ++k) {
                const std::complex<double> wk = std::polar(1.0, ang * k);
                const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * wk;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

}  // namespace

// Rows [r0, r1) of the line operator whose filter has the circular
// convolution kernel c: D[i][k] = sum over the window/reflection terms (q, a)
// of input k of c[(w + i - q) mod F] * a.
static std::vector<double> line_operator(const SegmentShape &s, const std::vector<double> &c, int r0, int r1) {
    const int F = s.F, w = s.w, nm = s.n_main, K = s.K(), vel = s.field, nr = r1 - r0;
    const std::vector<double> win = window_coefficients(w, s.patch_error);
    std::vector<double> D((size_t)nr * K, 0.0);
    auto add = [&](int k, int q, double a) {
        for (int i = 0; i < nr; ++i) D[(size_t)i * K + k] += c[((w + r0 + i - q) % F + F) % F] * a;
    };
    for (int k = 0; k < K; ++k) {
        if (k < w) {
            add(k, k, s.c_s1 * win[k]);  // dom1[k] <- side1 tail
        } else if (k < w + nm) {
            const int e = k - w;
            add(k, w + e, 1.0);  // the main line itself
            const int ql = vel ? w - e : w - 1 - e;  // dom1 position that reflects main[e]
            if (ql >= 0 && ql < w) add(k, ql, s.c_m1 * win[ql]);
            const int jr = vel ? nm - 2 - e : nm - 1 - e;  // dom3 position that reflects main[e]
            if (jr >= 0 && jr < w) add(k, w + nm + jr, s.c_m2 * win[w + 1 + jr]);
        } else {
            const int j = k - w - nm;
            add(k, w + nm + j, s.c_s2 * win[w + 1 + j]);  // dom3[j] <- side2 head
        }
    }
    return D;
}

static std::vector<double> convolution_kernel(std::vector<std::complex<double>> h) {
    idft_pow2(h);
    std::vector<double> c(h.size());
    for (size_t r = 0; r < h.size(); ++r) c[r] = h[r].real();
    return c;
}

std::vector<double> derivative_matrix(const SegmentShape &s) {
    if (!s.valid()) throw std::invalid_argument("derivative_matrix: inconsistent segment shape");
    // c = circular convolution kernel of the filter (1/F already in H)
    return line_operator(s, convolution_kernel(derivative_filter(s.F, s.table_len, s.field, s.dx)), 0, s.n_out());
}

std::vector<double> operator_row(const SegmentShape &s, const std::vector<std::complex<double>> &H, int row) {
    if (!s.valid() || (int)H.size() != s.F || row < 0 || row >= s.n_out())
        throw std::invalid_argument("operator_row: inconsistent segment shape, filter or row");
    return line_operator(s, convolution_kernel(H), row, row + 1);
}

}  // namespace pstdamd
