// Drives the C++ host mirror (crava_amd/host) the way the reference's
// Crava::Crava / computePostMeanResidAndFFTCov sequence drives FFTGrid and
// SeismicParametersHolder, on inputs written by tests/test_gpu_host_mirror.py,
// and writes the outputs for that test to compare with the oracle.
//
//   host_mirror_main <case dir>
//
// Inputs in <dir>: params.txt, wavelets.bin, corrxy.bin, corrt.bin,
// mu{0,1,2}.bin, data{l}.bin (raw float32).  Outputs: circ.bin,
// out_mu{c}.bin, out_res{l}.bin (time domain), out_cov{q}.bin (transformed),
// out_sim{c}.bin, out_scalars.txt.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../crava_amd/host/crava_host.hpp"

using crava_hip::FFTGrid;

static std::vector<float> read_bin(const std::string &path, size_t n) {
  std::vector<float> v(n);
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f || std::fread(v.data(), sizeof(float), n, f) != n) throw std::runtime_error("cannot read " + path);
  std::fclose(f);
  return v;
}

static void write_bin(const std::string &path, const std::vector<float> &v) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(v.data(), sizeof(float), v.size(), f) != v.size()) throw std::runtime_error("cannot write " + path);
  std::fclose(f);
}

// Read a real grid with the cursor, as the reference's writers do.
static std::vector<float> drain_real(FFTGrid *g) {
  g->setAccessMode(FFTGrid::READ);
  std::vector<float> v((size_t)g->getrsize());
  for (auto &x : v) x = g->getNextReal();
  g->endAccess();
  return v;
}

static std::vector<float> drain_complex(FFTGrid *g) {
  g->setAccessMode(FFTGrid::READ);
  std::vector<float> v(2 * (size_t)g->getcsize());
  for (size_t i = 0; i < v.size() / 2; ++i) {
    const crava_hip::fftw_complex_f c = g->getNextComplex();
    v[2 * i] = c.re;
    v[2 * i + 1] = c.im;
  }
  g->endAccess();
  return v;
}

static void fill_real(FFTGrid *g, const std::vector<float> &v) {
  g->setAccessMode(FFTGrid::WRITE);
  for (float x : v) g->setNextReal(x);
  g->endAccess();
}

int main(int argc, char **argv) {
  if (argc != 2) {
    std::cerr << "usage: host_mirror_main <case dir>\n";
    return 2;
  }
  const std::string dir = argv[1];
  try {
    std::ifstream in(dir + "/params.txt");
    crava_hip::Crava::Settings s;
    int ncorrt = 0, lowIntCut = 0;
    double angRange = 10.0;
    unsigned long long seed = 0;
    in >> s.nx >> s.ny >> s.nz >> s.nxp >> s.nyp >> s.nzp >> s.ntheta >> s.dz >> s.lowCut >> s.highCut >> s.wnc >>
        s.precision >> s.mixedThreshold >> s.corrGradI >> s.corrGradJ >> lowIntCut >> ncorrt >> angRange >> seed;
    const int nt = s.ntheta;
    s.thetaDeg.resize(nt);
    s.snr.resize(nt);
    s.A.resize(3 * nt);
    s.waveletNorm.resize(nt);
    double sigma0[3][3];
    for (auto &x : s.thetaDeg) in >> x;
    for (auto &x : s.snr) in >> x;
    for (auto &x : s.A) in >> x;
    for (auto &x : s.waveletNorm) in >> x;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) in >> sigma0[a][b];
    if (!in) throw std::runtime_error("bad params.txt");
    s.wavelets = read_bin(dir + "/wavelets.bin", (size_t)nt * s.nzp);
    s.corrXY = read_bin(dir + "/corrxy.bin", (size_t)s.nxp * s.nyp);
    const std::vector<float> corrT = read_bin(dir + "/corrt.bin", (size_t)ncorrt);
    const size_t rsize = (size_t)2 * (s.nxp / 2 + 1) * s.nyp * s.nzp;

    crava_hip::Context ctx;
    crava_hip::SeismicParametersHolder sph(&ctx);
    sph.allocateGrids(s.nx, s.ny, s.nz, s.nxp, s.nyp, s.nzp);
    sph.setPriorVar0(sigma0);
    if (s.precision == CRAVA_HIP_MIXED) sph.setCovTransformPrecision(CRAVA_HIP_MIXED);
    sph.initializeCorrelations(s.corrXY, corrT, s.corrGradI, s.corrGradJ, lowIntCut, s.nzp);
    write_bin(dir + "/circ.bin", sph.computeCircCorrT(corrT, lowIntCut, s.nzp));
    for (int c = 0; c < 3; ++c) fill_real(sph.mu(c), read_bin(dir + "/mu" + std::to_string(c) + ".bin", rsize));

    std::vector<FFTGrid *> seis;
    for (int l = 0; l < nt; ++l) {
      FFTGrid *g = new FFTGrid(&ctx, s.nx, s.ny, s.nz, s.nxp, s.nyp, s.nzp);
      g->setType(FFTGrid::DATA);
      g->createRealGrid();
      fill_real(g, read_bin(dir + "/data" + std::to_string(l) + ".bin", rsize));
      seis.push_back(g);
    }

    // local-noise scales (used by correctAlphaBetaRho below)
    s.localNoiseScale.assign(nt, std::vector<double>((size_t)s.nx * s.ny));
    for (int a = 0; a < nt; ++a)
      for (int j = 0; j < s.ny; ++j)
        for (int i = 0; i < s.nx; ++i) s.localNoiseScale[a][(size_t)j * s.nx + i] = 1.0 + ((i * 7 + j * 3 + a * 5) % 11) / 5.0;
    crava_hip::Crava crava(&ctx, s, seis);
    crava.computeDataVariance();
    crava.setupErrorCorrelation([angRange](float dTheta) { return std::exp(-3.0 * std::fabs(dTheta) / angRange); });
    const long long nfail = crava.computePostMeanResidAndFFTCov(sph);

    for (int c = 0; c < 3; ++c) write_bin(dir + "/out_mu" + std::to_string(c) + ".bin", drain_real(sph.mu(c)));
    for (int l = 0; l < nt; ++l) write_bin(dir + "/out_res" + std::to_string(l) + ".bin", drain_real(seis[l]));
    for (int q = 0; q < 6; ++q) write_bin(dir + "/out_cov" + std::to_string(q) + ".bin", drain_complex(sph.cov(q)));

    // one posterior realisation (posterior mean added in the time domain)
    FFTGrid *real[3];
    for (int c = 0; c < 3; ++c) {
      real[c] = new FFTGrid(&ctx, s.nx, s.ny, s.nz, s.nxp, s.nyp, s.nzp);
      real[c]->setType(FFTGrid::PARAMETER);
      real[c]->createComplexGrid();
    }
    const long long simfail = crava.simulate(sph, seed, 0, real);
    for (int c = 0; c < 3; ++c) write_bin(dir + "/out_sim" + std::to_string(c) + ".bin", drain_real(real[c]));
    // the same with the reference's RandomGen stream from state 23665
    uint32_t rg = 23665;
    crava.simulate(sph, rg, real);
    for (int c = 0; c < 3; ++c) write_bin(dir + "/out_simref" + std::to_string(c) + ".bin", drain_real(real[c]));
    std::ofstream(dir + "/simref_seed.txt") << rg << "\n";

    // local-noise correction (host computeG) of the posterior means in
    // place, towards the prior means, as the reference corrects postAlpha_
    // etc.: postVar0 = D sigma0 D with D = diag(0.6, 0.7, 0.8) (distinct
    // eigenvalues of sigma0^-1/2 (sigma0 - postVar0) sigma0^-1/2, so that its
    // eigenvectors are defined up to sign); then the same reference-stream
    // realisation, now with the local-noise factors and the corrected means
    // added (src/crava.cpp:1547-1568)
    {
      const double dg[3] = {0.6, 0.7, 0.8};
      double smd[3][3];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) smd[a][b] = (dg[a] * sigma0[a][b]) * dg[b];
      FFTGrid *post[3], *mean2[3];
      std::vector<std::vector<float>> uncorrected(3);
      for (int c = 0; c < 3; ++c) {
        uncorrected[c] = drain_real(sph.mu(c));
        post[c] = sph.mu(c);
        mean2[c] = new FFTGrid(&ctx, s.nx, s.ny, s.nz, s.nxp, s.nyp, s.nzp);
        mean2[c]->setType(FFTGrid::PARAMETER);
        mean2[c]->createRealGrid();
        fill_real(mean2[c], read_bin(dir + "/mu" + std::to_string(c) + ".bin", rsize));
      }
      crava.correctAlphaBetaRho(sph.getPriorVar0(), smd, post, mean2);
      for (int c = 0; c < 3; ++c) {
        write_bin(dir + "/out_lnc" + std::to_string(c) + ".bin", drain_real(sph.mu(c)));
        delete mean2[c];
      }
      uint32_t rg2 = 23665;
      crava.simulate(sph, rg2, real);
      for (int c = 0; c < 3; ++c) write_bin(dir + "/out_simlnc" + std::to_string(c) + ".bin", drain_real(real[c]));
      // the steps below work on the uncorrected posterior means
      for (int c = 0; c < 3; ++c) fill_real(sph.mu(c), uncorrected[c]);
    }
    for (int c = 0; c < 3; ++c) delete real[c];

    // forward modelling of the posterior means (Crava::computeSyntSeismic),
    // with a stretch table on angle 0 and a local time shift on the last
    const size_t ntr = (size_t)s.nx * s.ny;
    std::vector<float> sf(ntr), ts(ntr);
    for (int j = 0; j < s.ny; ++j)
      for (int i = 0; i < s.nx; ++i) {
        sf[(size_t)j * s.nx + i] = 0.9f + 0.2f * (float)((i * 7 + j * 3) % 11) / 10.0f;
        ts[(size_t)j * s.nx + i] = (float)((i + 2 * j) % 9) - 4.0f;
      }
    s.stretch.assign(nt, {});
    s.localShift.assign(nt, {});
    s.stretch[0] = sf;
    s.localShift[nt - 1] = ts;
    write_bin(dir + "/stretch.bin", sf);
    write_bin(dir + "/shift.bin", ts);
    crava_hip::Crava fwd(&ctx, s, seis);
    std::vector<FFTGrid *> syn(nt);
    for (int l = 0; l < nt; ++l) {
      syn[l] = new FFTGrid(&ctx, s.nx, s.ny, s.nz, s.nxp, s.nyp, s.nzp);
      syn[l]->setType(FFTGrid::DATA);
      syn[l]->createRealGrid();
    }
    fwd.computeSyntSeismic(sph.mu(0), sph.mu(1), sph.mu(2), syn);
    for (int l = 0; l < nt; ++l) {
      write_bin(dir + "/out_syn" + std::to_string(l) + ".bin", drain_real(syn[l]));
    }

    // non-constant thickness (Crava::divideDataByScaleWavelet with
    // interpolateSeismic) on the synthetic seismic, two traces of each angle
    // emptied first, with a prior holder whose covariance grids are still in
    // the time domain, as in the reference's Crava constructor
    {
      crava_hip::SeismicParametersHolder sph2(&ctx);
      sph2.allocateGrids(s.nx, s.ny, s.nz, s.nxp, s.nyp, s.nzp);
      sph2.setPriorVar0(sigma0);
      sph2.initializeCorrelations(s.corrXY, corrT, s.corrGradI, s.corrGradJ, lowIntCut, s.nzp);
      crava_hip::Crava::Settings s2 = s;
      s2.stretch.clear();
      s2.relThick.resize((size_t)s.nxp * s.nyp);
      for (int j = 0; j < s.nyp; ++j)
        for (int i = 0; i < s.nxp; ++i)
          s2.relThick[(size_t)j * s.nxp + i] = 0.7 + 0.3 * (double)((i * 5 + j * 3) % 13) / 12.0;
      std::vector<float> gain(ntr);
      for (size_t t = 0; t < ntr; ++t) gain[t] = 0.8f + 0.05f * (float)(t % 9);
      s2.localGain.assign(nt, {});
      s2.localGain[0] = gain;
      s2.energyThreshold = 0.05f;
      for (int l = 0; l < nt; ++l) {
        syn[l]->setAccessMode(FFTGrid::RANDOMACCESS);
        for (int k = 0; k < s.nzp; ++k) {
          syn[l]->setRealValue(2, 1, k, 0.0f, true);
          syn[l]->setRealValue(3, 1, k, 0.0f, true);
        }
        syn[l]->endAccess();
      }
      // energy matching on angle 0 (Crava::computeVariances, in the
      // constructor before the division)
      s2.matchEnergies.assign(nt, 0);
      s2.matchEnergies[0] = 1;
      crava_hip::Crava cnt(&ctx, s2, syn);
      const std::vector<float> corrTv = sph2.extractParamCorrFromCovAlpha(s.nzp);
      cnt.computeVariances(corrTv, sph2,
                           [angRange](float dTheta) { return std::exp(-3.0 * std::fabs(dTheta) / angRange); });
      write_bin(dir + "/nc_corrt.bin", corrTv);
      write_bin(dir + "/nc_errvar.bin", cnt.getErrorVariance());
      write_bin(dir + "/nc_wavelets.bin", cnt.getWavelets());
      {
        std::vector<float> e;
        for (const auto *v : {&cnt.getModelVariance(), &cnt.getSignalVariance(), &cnt.getTheoSNRatio()})
          e.insert(e.end(), v->begin(), v->end());
        write_bin(dir + "/nc_energy.bin", e);
      }
      cnt.divideDataByScaleWavelet(sph2);
      std::vector<float> ev(nt);
      for (int l = 0; l < nt; ++l) ev[l] = (float)cnt.getErrThetaCov()[(size_t)l * nt + l];
      write_bin(dir + "/swd_errvar.bin", ev);
      write_bin(dir + "/swd_gain.bin", gain);
      write_bin(dir + "/swd_corrt.bin", sph2.getPriorCorrTFiltered(s.nz, s.nzp));
      for (int l = 0; l < nt; ++l) write_bin(dir + "/out_swd" + std::to_string(l) + ".bin", drain_real(syn[l]));
      // the inversion of the divided data with the non-constant-thickness
      // per-k vectors (K = D A), prior means from the inputs
      for (int c = 0; c < 3; ++c) fill_real(sph2.mu(c), read_bin(dir + "/mu" + std::to_string(c) + ".bin", rsize));
      const long long ncfail = cnt.computePostMeanResidAndFFTCov(sph2);
      for (int c = 0; c < 3; ++c) write_bin(dir + "/nc_mu" + std::to_string(c) + ".bin", drain_real(sph2.mu(c)));
      for (int l = 0; l < nt; ++l) write_bin(dir + "/nc_res" + std::to_string(l) + ".bin", drain_real(syn[l]));
      for (int q = 0; q < 6; ++q) write_bin(dir + "/nc_cov" + std::to_string(q) + ".bin", drain_complex(sph2.cov(q)));
      std::ofstream(dir + "/nc_fail.txt") << ncfail << "\n";
    }
    for (int l = 0; l < nt; ++l) delete syn[l];

    // CRAVA binary files (FFTGrid::writeCravaFile / readCravaFile): write the
    // posterior mean of alpha, read the file the test wrote, and report a
    // grid of other dimensions
    crava_hip::FFTGrid::CravaHeader hdr;
    hdr.x0 = 456000.5;
    hdr.y0 = 6780000.25;
    hdr.dx = 25.0;
    hdr.dy = 12.5;
    hdr.nx = s.nx;
    hdr.ny = s.ny;
    hdr.il0 = 1001.0;
    hdr.xl0 = 2002.0;
    hdr.ilStepX = 0.5;
    hdr.ilStepY = -0.25;
    hdr.xlStepX = 0.125;
    hdr.xlStepY = 2.0;
    hdr.angle = 0.3;
    sph.mu(0)->writeCravaFile(dir + "/mu0", hdr);
    {
      FFTGrid rb(&ctx, s.nx, s.ny, s.nz, s.nxp, s.nyp, s.nzp);
      std::string err;
      rb.readCravaFile(dir + "/in.crava", err);
      write_bin(dir + "/crava_read.bin", drain_real(&rb));
      FFTGrid other(&ctx, s.nx, s.ny, s.nz, s.nxp, s.nyp, s.nzp + 2);
      other.readCravaFile(dir + "/in.crava", err);
      other.readCravaFile(dir + "/missing.crava", err);
      std::ofstream(dir + "/crava_err.txt") << err;
    }

    // STORM files of the posterior mean of alpha: binary with surface names,
    // ascii flat, and the padded binary cube
    {
      crava_hip::FFTGrid::StormHeader sh;
      sh.xmin = 456000.5;
      sh.lx = 11200.0;
      sh.ymin = 6780000.25;
      sh.ly = 5600.0;
      sh.lz = 1024.0;
      sh.angle = 0.3;
      sh.topName = "top_surface";
      sh.botName = "base_surface";
      sph.mu(0)->writeStormFile(dir + "/mu0", sh, false);
      sph.mu(0)->writeStormFile(dir + "/mu0_flat", sh, true, false, true);
      sph.mu(0)->writeStormFile(dir + "/mu0_pad", sh, false, true);
      std::ofstream(dir + "/storm_type.txt") << sph.mu(0)->getType() << "\n";
    }

    // SEG-Y (every trace header format) and SGRI files of the posterior
    // mean of alpha on a rotated simbox with uneven top and base surfaces,
    // one trace without a top; the geometry goes to segy_geom.txt / .bin
    {
      crava_hip::FFTGrid::SimboxGeometry g;
      g.xmin = 456000.5;
      g.ymin = 6780000.25;
      g.dx = 25.0;
      g.dy = 12.5;
      g.angle = 0.3;
      g.dz = s.dz;
      g.lz = s.nz * s.dz;
      g.il0 = 1001.0;
      g.xl0 = 2002.0;
      g.ilStepX = 0.0348;
      g.ilStepY = -0.0107;
      g.xlStepX = 0.0214;
      g.xlStepY = 0.0693;
      g.top.resize(ntr);
      g.bot.resize(ntr);
      double zmin = 1e30, zmax = -1e30;
      for (int j = 0; j < s.ny; ++j)
        for (int i = 0; i < s.nx; ++i) {
          const size_t t = (size_t)j * s.nx + i;
          g.top[t] = 1800.0 + 3.7 * i - 2.1 * j + 0.37 * ((i * 7 + j * 5) % 11);
          g.bot[t] = g.top[t] + g.lz * (0.75 + 0.25 * (double)((i * 3 + j * 2) % 7) / 6.0);
          zmin = std::min(zmin, g.top[t]);
          zmax = std::max(zmax, g.bot[t]);
        }
      g.top[(size_t)1 * s.nx + 2] = -99999.0;   // a trace with no top surface
      g.zmin = zmin;
      g.zmax = zmax;
      const float z0 = (float)(std::floor(zmin / 10.0) * 10.0);
      for (int f = 0; f < 5; ++f) sph.mu(0)->writeSegyFile(dir + "/mu0_thf" + std::to_string(f), g, z0, f);
      sph.mu(0)->writeSgriFile(dir + "/mu0", g, "Posterior alpha");
      std::ofstream gt(dir + "/segy_geom.txt");
      gt.precision(17);
      gt << g.xmin << " " << g.ymin << " " << g.dx << " " << g.dy << " " << g.angle << " " << g.lz << " " << g.dz << " "
         << g.zmin << " " << g.zmax << " " << g.il0 << " " << g.xl0 << " " << g.ilStepX << " " << g.ilStepY << " "
         << g.xlStepX << " " << g.xlStepY << " " << z0 << "\n";
      std::ofstream gb(dir + "/segy_geom.bin", std::ios::binary);
      gb.write(reinterpret_cast<const char *>(g.top.data()), (std::streamsize)(ntr * sizeof(double)));
      gb.write(reinterpret_cast<const char *>(g.bot.data()), (std::streamsize)(ntr * sizeof(double)));
    }

    // time-lapse merge (State4D::merge) when the test supplies the parts:
    // 3 + 3 means and 6 + 6 + 9 covariance grids, transformed, in that order
    {
      std::ifstream probe(dir + "/state4d_in.bin", std::ios::binary);
      if (probe) {
        const size_t csz = 2 * (size_t)(s.nxp / 2 + 1) * s.nyp * s.nzp;
        const std::vector<float> all = read_bin(dir + "/state4d_in.bin", 27 * csz);
        std::vector<FFTGrid *> parts(27);
        for (int g = 0; g < 27; ++g) {
          parts[g] = new FFTGrid(&ctx, s.nx, s.ny, s.nz, s.nxp, s.nyp, s.nzp);
          parts[g]->createComplexGrid();
          parts[g]->setAccessMode(FFTGrid::WRITE);
          for (size_t i = 0; i < csz / 2; ++i)
            parts[g]->setNextComplex(crava_hip::fftw_complex_f{all[g * csz + 2 * i], all[g * csz + 2 * i + 1]});
          parts[g]->endAccess();
        }
        auto range = [&](int a, int b) { return std::vector<FFTGrid *>(parts.begin() + a, parts.begin() + b); };
        crava_hip::State4D st(&ctx, range(0, 3), range(3, 6), range(6, 12), range(12, 18), range(18, 27));
        st.merge(sph);
        std::vector<float> out;
        for (int c = 0; c < 3; ++c) {
          const std::vector<float> v = drain_complex(sph.mu(c));
          out.insert(out.end(), v.begin(), v.end());
        }
        FFTGrid *sig[6] = {sph.GetCovAlpha(), sph.GetCrCovAlphaBeta(), sph.GetCrCovAlphaRho(), sph.GetCovBeta(),
                           sph.GetCrCovBetaRho(), sph.GetCovRho()};
        for (FFTGrid *g : sig) {
          const std::vector<float> v = drain_complex(g);
          out.insert(out.end(), v.begin(), v.end());
        }
        write_bin(dir + "/state4d_out.bin", out);
        for (FFTGrid *g : parts) delete g;
      }
    }

    // SeismicParametersHolder::createPostCov00 / getOrigin / updatePriorVar
    // on the covariance grids taken back to the time domain
    sph.invFFTCovGrids();
    write_bin(dir + "/postcov00.bin", sph.createPostCov00(sph.cov(0)));
    {
      std::vector<float> org(6), pv(9);
      for (int q = 0; q < 6; ++q) {
        org[q] = sph.getOrigin(sph.cov(q));
        write_bin(dir + "/out_covt" + std::to_string(q) + ".bin", drain_real(sph.cov(q)));
      }
      sph.updatePriorVar();
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) pv[3 * a + b] = (float)sph.getPriorVar0()[a][b];
      write_bin(dir + "/origins.bin", org);
      write_bin(dir + "/priorvar.bin", pv);
    }

    // an ill-conditioned prior covariance is rejected (the reference's
    // robustification is not reproducible); a well-conditioned one is kept
    {
      const double bad[3][3] = {{1.0, 0.9999, 0.0}, {0.9999, 1.0, 0.0}, {0.0, 0.0, 1.0}};
      bool rejected = false;
      try {
        sph.setPriorVar0(bad);
      } catch (const crava_hip::Error &) {
        rejected = true;
      }
      sph.setPriorVar0(sigma0);
      std::ofstream(dir + "/illcond.txt") << (rejected ? 1 : 0) << "\n";
    }

    // a few cursor / random-access checks against the values just written
    FFTGrid *m0 = sph.mu(0);
    m0->setAccessMode(FFTGrid::RANDOMACCESS);
    const float v = m0->getRealValue(1, 2, 3);
    const float outside = m0->getRealValue(s.nx, 0, 0);
    m0->setRealValue(1, 2, 3, v);
    m0->endAccess();

    std::ofstream out(dir + "/out_scalars.txt");
    out.precision(9);
    out << nfail << " " << simfail << " " << crava.getFP64Points() << "\n";
    for (float x : crava.getDataVariance()) out << x << " ";
    out << "\n";
    for (float x : crava.getErrorVariance()) out << x << " ";
    out << "\n";
    out.precision(17);
    for (double x : crava.getErrThetaCov()) out << x << " ";
    out.precision(9);
    out << "\n" << v << " " << outside << "\n";
    for (FFTGrid *g : seis) delete g;
  } catch (const crava_hip::Error &e) {
    std::cerr << "crava_hip error " << e.status() << ": " << e.what() << "\n";
    return 1;
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
