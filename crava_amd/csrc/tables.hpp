// Host-side per-k preparation of the posterior loop (constant-thickness case).
#pragma once
#include <cstdint>

namespace crava {

// Per-k vectors of Crava::computePostMeanResidAndFFTCov (src/crava.cpp:1042-1057):
//   invert[k] : 5*minRelThick < f_k < 55-style frequency gate (:1044,1098)
//   kWD[k][l] : kD(k) * conj(W_l(k))            (:1050-1051, fillkW :1337-1345)
//   e1[k][l]  : conj(W_l(k)) / |w_l|            (fillkWNorm :1347-1356)
//   e2[k][l]  : conj(E3_l(k)) / |e3_l|, e3 = central(backward(forward(w_l)))
//   e3[k][l]  : 1                                (:1057)
// Spectra are 1-D DFTs with FFTW's forward sign, computed in double and
// rounded to float; k >= nzp/2+1 uses the Hermitian extension of
// Wavelet::getCAmp (src/wavelet.cpp:309-323).
void perk_vectors(int ntheta, int nz, int nzp, double lz, float low_cut, float high_cut, float min_rel_thick,
                  const float *wavelets, const float *wnorm, unsigned char *invert, float *kWD, float *e1,
                  float *e2, float *e3);

// The same per-k vectors for a simbox that is not of constant thickness
// (src/crava.cpp:1058-1083; the data were divided by the wavelets first):
//   kWD[k][l] : kD(k)                            (K = D A)
//   e1[k][l]  : 1 / |w_l|
//   e2[k][l]  : kD3(k) / |e3_l|, kD3 the spectrum of central(backward(d))
//   e3[k][l]  : fillInverseAbskWRobust (:1361-1384): 1/sqrt(max(|W_l(k)|^2,
//               |W_l(k, 0.999)|^2)), or |w_l| nzp^2 100 where both vanish.
void perk_vectors_nonconst(int ntheta, int nz, int nzp, double lz, float low_cut, float high_cut,
                           float min_rel_thick, const float *wavelets, const float *wnorm, unsigned char *invert,
                           float *kWD, float *e1, float *e2, float *e3);

// Wavelet::resample (src/wavelet.cpp:416-463): nzp samples in FFT order at
// spacing dz from a wavelet of nw samples at spacing dzw centred on sample
// `center`, each by the six-sample Ricker-correlation kriging of
// Wavelet::getWaveletValue; *norm receives Wavelet::norm_ (sqrt of the sum
// of squares).  Returns false where a kriging matrix is not positive definite
// (NRLib::CholeskySolve throws there).
bool wavelet_resample(const float *w, int nw, int center, float dzw, float dz, int nzp, float *out, float *norm);

// Crava::computeVariances' model variance and energy-matching gain
// (src/crava.cpp:400-466, computeWDCorrMVar :1763-1775).
void wavelet_energy(int ntheta, int nzp, const float *wavelets, const float *A, const double *prior_var0,
                    const float *corr_t, const float *error_variance, const float *emp_sn,
                    const unsigned char *match, float *gain, float *model_var, float *signal_var,
                    float *theo_sn);

// FFTGrid::findClosestFactorableNumber (src/fftgrid.cpp:2837-2866): the
// smallest 2^a 3^b 5^c 7^d >= leastint within the reference's loop bounds
// (ceil of float logarithms), else 2^ceil(log2 leastint).
int closest_factorable(int leastint);
// ModelGeneral::setPaddingSize (src/modelgeneral.cpp:581-587)
int padding_size(int n, double pad_fac);
// ModelGeneral::estimateXYPaddingSizes / estimateZPaddingSize
// (src/modelgeneral.cpp:1337-1425) without the logging.
void estimate_xy_padding(int nx, int ny, double dx, double dy, double lx, double ly, double x_pad_fac,
                         double y_pad_fac, bool estimate, float range1, float range2, float angle, int *nxpad,
                         int *nypad, double *true_x_pad_fac, double *true_y_pad_fac);
void estimate_z_padding(int nz, double lz, double min_rel_thick, double z_pad_fac, bool estimate,
                        float default_wavelet_length, int *nzpad, double *true_z_pad_fac);

// SeismicParametersHolder::setCorrelationParameters' robustification of
// Sigma0 (src/seismicparametersholder.cpp:114-135) with DGEEV's eigenvalues
// and eigenvectors (dgeev3; a diagonal Sigma0 in closed form): returns 0
// when every eigenvalue is at least 0.1 % of the largest (out = in), 1 when
// out was lifted, -1 when it would be lifted but DGEEV's eigenvectors are
// not reproduced here (a zero off-diagonal entry or balancing).
int prior_var0_robustify(const double in[9], double out[9]);
// computeCircCorrT: nzp floats into out; false where the reference exits
// (c(0) <= 1e-5).
bool circ_corr_t(const float *prior, int n, int min_int_fq, int nzp, float *out);

// Half spectrum (nzp/2+1 {re, im} pairs) of a real series of nzp samples
// with FFTW's forward sign and no normalisation (Wavelet::fft1DInPlace,
// src/wavelet.cpp:246-261): double accumulation rounded to float, DC and
// Nyquist bins exactly real.
void wavelet_half_spectrum(const float *w, int nzp, float *re_im);

// FFTGrid::fillInComplexNoise (src/fftgrid.cpp:1214-1263) with the
// reference's RandomGen stream, on the host: nxp/2+1 x nyp x nzp {re, im}
// pairs; *seed is the generator state, advanced as the reference's is.
void reference_complex_noise(int nxp, int nyp, int nzp, uint32_t *seed, float *re_im);

}  // namespace crava
