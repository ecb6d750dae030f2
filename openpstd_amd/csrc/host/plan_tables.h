// Device plan cache tables of the MI355X solver, assembled in double on the host:
// the spectral derivative filter of a line, the window, and the dense form of
// one segment's windowed line derivative (the matrix the GEMM path multiplies).
#pragma once

#include <complex>
#include <vector>

namespace pstdamd {

// Derivative filter H[k], k < F, of a line of FFT length F: derfact of
// WisdomCache::discretize_wave_numbers (kernel/core/WisdomCache.cpp:53-75) for a
// wavenumber table of length L (k_m = m * (pi/dx) / (L/2); pressure
// exp(-i k dx/2) i k, velocity exp(+i k dx/2) i k), on bins 0..F/2 as
// spatderp3 applies it (kernel/core/kernel_functions.cpp:234, 294), bins 0 and
// F/2 real (FFTW's c2r ignores their imaginary parts), extended
// Hermitian-symmetrically to F bins, 1/F (kernel_functions.cpp:243, 306) folded
// in.  field: 0 pressure, 1 velocity.
std::vector<std::complex<double>> derivative_filter(int F, int L, int field, double dx);

// Fractional-shift filter of a line for spectral interpolation: the factors
// Receiver::get_fft_factors forms (kernel/core/Receiver.cpp:57-73),
// exp(i k_m delta dx) with k_m as in derivative_filter (so exp(i pi m delta /
// (L/2))) on bins 0..F/2, bin F/2 real, Hermitian-extended, 1/F folded in.
// Applied to a pressure line in place of derfact, output i is the
// band-limited line at position i + delta (cells).
std::vector<std::complex<double>> shift_filter(int F, int L, double delta);

// get_window_coefficients (kernel/core/kernel_functions.cpp:327-333): 2w+1 values.
std::vector<double> window_coefficients(int w, double patch_error);

// Line geometry of one segment (one spatderp3 call of Domain::calc).
struct SegmentShape {
    int F;            // FFT length
    int table_len;    // wavenumber table length (>= F)
    int field;        // 0 pressure (n_main = n, n_out = n+1), 1 velocity (n_main = n+1, n_out = n)
    double dx;
    int w;            // window half length
    int n_main;       // main values of a line
    double patch_error;
    double c_s1, c_m1, c_s2, c_m2;  // get_rho_array stencil coefficients
    int n_out() const { return field == 0 ? n_main + 1 : n_main - 1; }
    int K() const { return n_main + 2 * w; }
    bool valid() const {
        return F > 0 && (F & (F - 1)) == 0 && table_len >= F && w > 0 && n_main >= w + 1 && n_main + 2 * w <= F &&
               (field == 0 || field == 1) && dx > 0;
    }
};

// Dense derivative matrix D (n_out x K, row-major) of a segment: for the raw
// input line u = [side1 tail (w values) | main (n_main) | side2 head (w)],
// D u equals spatderp3's output for that line (the windowed line
// [dom1 | main | dom3 | 0] transformed, multiplied by the filter, transformed
// back and cropped, kernel/core/kernel_functions.cpp:191-307).  The window and
// reflections make the line a linear map W (F x K, at most three terms per
// input) and the filter a circular convolution with c = IDFT(H), so
// D[i][k] = sum over W's terms (q, a) of input k of c[(w + i - q) mod F] * a.
// Side tails are read from the neighbour's last w (pressure) or its values
// w+1 .. 2 from the end (velocity), heads from its first w (pressure) or
// values 1..w (velocity); the caller supplies them in that order.
std::vector<double> derivative_matrix(const SegmentShape &s);

// Row `row` (< n_out) of the same operator with the filter H (F bins, 1/F
// folded in) in place of the derivative filter: K = n_main + 2w weights of
// [side1 tail | main | side2 head].
std::vector<double> operator_row(const SegmentShape &s, const std::vector<std::complex<double>> &H, int row);

}  // namespace pstdamd
