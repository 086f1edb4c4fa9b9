// fft.cpp -- the reference's monitor-series transform (src/data_utils.cpp:325-427).
//
// fft() keeps the first T = 2^floor(log2 N) samples but uses the phase step 2 pi / N of the
// untruncated length, so for N not a power of two it is not a DFT of either length.  The
// recursion below visits index pairs in the reference's order (split by stride, pair x[rem]
// with x[rem + T/2] at the leaves) and accumulates every output bin in that order, so the
// result matches the reference to the last bit; the closed form it computes is
//   X[k mod T] = sum_{n < T} x[n] exp(-i 2 pi n k / N),  k = -T/2 .. T/2 - 1.
#include <cmath>
#include <complex>
#include <vector>

#include "sjf_host.hpp"

namespace sjfh {

namespace {

struct Cx {
    double re, im;
};
Cx cexp_i(double im) { return {std::cos(im), std::sin(im)}; }
Cx cmul(Cx z, Cx w) { return {z.re * w.re - z.im * w.im, z.re * w.im + z.im * w.re}; }

struct Helper {
    const std::vector<Cx>& x;
    unsigned long long T;
    double tau;
    std::vector<Cx>& out;
};

void part(const Helper& h, unsigned long long span, unsigned long long rem) {
    if (span * 2 >= h.T) {
        const long half = (long)(h.out.size() / 2);
        unsigned long long k_out = (unsigned long long)half;
        for (long k = -half; k < half; ++k) {
            Cx v = cmul(cexp_i(-h.tau * span * k), h.x[span + rem]);
            v.re += h.x[rem].re;
            v.im += h.x[rem].im;
            v = cmul(cexp_i(-h.tau * rem * k), v);
            h.out[k_out].re += v.re;
            h.out[k_out].im += v.im;
            if (++k_out == h.out.size()) k_out = 0;
        }
    } else {
        part(h, 2 * span, rem);
        part(h, 2 * span, rem + span);
    }
}

}  // namespace

std::vector<std::complex<double>> fft(const std::vector<std::complex<double>>& in) {
    if (in.empty()) return {};
    const unsigned long long log2n = (unsigned long long)(std::log((double)in.size()) / std::log(2));
    const unsigned long long T = 1ull << log2n;
    std::vector<Cx> x(in.size());
    for (size_t i = 0; i < in.size(); ++i) x[i] = {in[i].real(), in[i].imag()};
    std::vector<Cx> out(T, Cx{0, 0});
    Helper h{x, T, 2 * (double)M_PI / (double)in.size(), out};
    part(h, 1, 0);
    std::vector<std::complex<double>> r(T);
    for (size_t i = 0; i < T; ++i) r[i] = {out[i].re, out[i].im};
    return r;
}

}  // namespace sjfh
