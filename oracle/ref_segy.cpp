// ref_segy.cpp -- drives the reference's OWN NRLib SEG-Y writer
// (libs/nrlib/segy, iotools, the vendored boost subset; compiled from
// /root/reference by oracle/Makefile into oracle/_ref/libcrava_ref_segy.so)
// in the order FFTGrid::writeSegyFile uses it, and restates
// FFTGrid::writeSgriFile and FFTGrid::writeStormFile (with
// Simbox::getStormHeader) on NRLib's ToString and binary float writer.
//
// TEST INFRASTRUCTURE ONLY (the checker): the product never links or loads
// it.  The Simbox and FFTGrid glue (trace centres, surfaces, the regular-z
// interpolation) is restated here because src/simbox.cpp and
// src/fftgrid.cpp need the rest of the reference.
#include <cmath>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "nrlib/iotools/fileio.hpp"
#include "nrlib/iotools/stringtools.hpp"
#include "nrlib/math/constants.hpp"
#include "nrlib/segy/commonheaders.hpp"
#include "nrlib/segy/segy.hpp"
#include "nrlib/segy/segygeometry.hpp"
#include "nrlib/segy/traceheader.hpp"

namespace {

const double kRMissing = -99999.000;    // src/definitions.h:35
const double kWellMissing = -999.00000; // src/definitions.h:39
#define RMISSING_STORM -99999.000       // RMISSING as src/definitions.h:35 defines it (a double literal)

struct Box {
  int nx, ny, nz, rnxp, nyp;
  const float *grid;                    // real padded [nzp][nyp][rnxp]
  double xmin, ymin, dx, dy, angle, lz, dz, zmin, zmax;
  const double *top, *bot;              // [ny][nx]
  // Simbox::getXYCoord, src/simbox.cpp:249-255
  void xy(int i, int j, double &x, double &y) const {
    const double rx = (static_cast<double>(i) + 0.5) * dx;
    const double ry = (static_cast<double>(j) + 0.5) * dy;
    x = rx * std::cos(angle) - ry * std::sin(angle) + xmin;
    y = rx * std::sin(angle) + ry * std::cos(angle) + ymin;
  }
  // Simbox::getRelThick, src/simbox.cpp:713-732
  double rel_thick(int i, int j) const {
    const double t = top[(size_t)j * nx + i], b = bot[(size_t)j * nx + i];
    return (t != kRMissing && b != kRMissing) ? (b - t) / lz : 1.0;
  }
  // FFTGrid::getRealValue (inside the volume, RMISSING outside)
  float value(int i, int j, int k) const {
    if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return static_cast<float>(kRMissing);
    return grid[(size_t)i + (size_t)rnxp * ((size_t)j + (size_t)nyp * k)];
  }
  // FFTGrid::getRegularZInterpolatedRealValue, src/fftgrid.cpp:2491-2522
  float regular_z(int i, int j, double z0Reg, double dzReg, int kReg, double z0Grid, double dzGrid) const {
    float z = static_cast<float>(z0Reg + dzReg * kReg);
    float t = static_cast<float>((z - z0Grid) / dzGrid);
    int kGrid = int(t);
    t -= kGrid;
    if (kGrid < 0) {
      if (kGrid < -1) return static_cast<float>(kRMissing);
      kGrid = 0;
      t = 0;
    } else if (kGrid > nz - 2) {
      if (kGrid > nz - 1) return static_cast<float>(kRMissing);
      kGrid = nz - 2;
      t = 1;
    }
    const float v1 = value(i, j, kGrid);
    const float v2 = value(i, j, kGrid + 1);
    return (1 - t) * v1 + t * v2;
  }
};

}  // namespace

extern "C" {

// FFTGrid::writeSegyFile, src/fftgrid.cpp:2189-2264; writes fileName + ".segy".
int ref_write_segy(const char *fileName, const float *grid, int nx, int ny, int nz, int rnxp, int nyp, double xmin,
                   double ymin, double dx, double dy, double angle, double lz, double dzs, double zmin, double zmax,
                   double il0, double xl0, double ilStepX, double ilStepY, double xlStepX, double xlStepY,
                   const double *top, const double *bot, float z0, int thfCode) {
  try {
    Box b{nx, ny, nz, rnxp, nyp, grid, xmin, ymin, dx, dy, angle, lz, dzs, zmin, zmax, top, bot};
    std::string gfName = std::string(fileName) + ".segy";
    NRLib::TextualHeader header = NRLib::TextualHeader::standardHeader();
    float dz = float(floor(dzs + 0.5));
    if (dz == 0) dz = 1.0;
    int segynz = int(ceil((zmax - z0) / dz));
    NRLib::TraceHeaderFormat thf(thfCode);
    NRLib::SegY *segy = new NRLib::SegY(gfName.c_str(), z0, segynz, dz, header, thf);
    NRLib::SegyGeometry *geometry =
        new NRLib::SegyGeometry(xmin, ymin, dx, dy, nx, ny, il0, xl0, ilStepX, ilStepY, xlStepX, xlStepY, angle);
    segy->SetGeometry(geometry);
    delete geometry;
    std::vector<float> trace(segynz);
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double x, y;
        b.xy(i, j, x, y);
        double z = top[(size_t)j * nx + i];
        if (z == kRMissing || z == kWellMissing) {
          for (int k = 0; k < segynz; k++) trace[k] = 0;
        } else {
          double gdz = dzs * b.rel_thick(i, j);
          int firstData = static_cast<int>(floor(0.5 + (z - z0) / dz));
          int endData = static_cast<int>(floor(0.5 + ((z - z0) + nz * gdz) / dz));
          if (endData > segynz) endData = segynz;
          int k;
          for (k = 0; k < firstData; k++) trace[k] = 0;
          for (; k < endData; k++) trace[k] = b.regular_z(i, j, z0, dz, k, z, gdz);
          for (; k < segynz; k++) trace[k] = 0;
          float xx = static_cast<float>(x);
          float yy = static_cast<float>(y);
          segy->StoreTrace(xx, yy, trace, NULL);
        }
      }
    segy->WriteAllTracesToFile();
    delete segy;
    return 0;
  } catch (...) {
    return 1;
  }
}

// FFTGrid::writeSgriFile, src/fftgrid.cpp:2325-2397; writes fileName +
// ".Sgrh" and ".Sgri".
int ref_write_sgri(const char *fileName, const char *label, const float *grid, int nx, int ny, int nz, int rnxp,
                   int nyp, double xmin, double ymin, double dx, double dy, double angle, double lz, double dzs,
                   double zmin, double zmax, const double *top, const double *bot) {
  try {
    Box b{nx, ny, nz, rnxp, nyp, grid, xmin, ymin, dx, dy, angle, lz, dzs, zmin, zmax, top, bot};
    double vertScale = 0.001;
    double horScale = 0.001;
    std::string fName = std::string(fileName) + ".Sgrh";
    std::ofstream headerFile;
    NRLib::OpenWrite(headerFile, fName);
    headerFile << "NORSAR General Grid Format v1.0\n";
    headerFile << "3\n";
    headerFile << "X (km)\n";
    headerFile << "Y (km)\n";
    headerFile << "T (s)\n";
    headerFile << "FFT-grid\n";
    headerFile << "1\n";
    headerFile << label << std::endl;
    headerFile << "1 1 1\n";
    double zMin = zmin, zMax = zmax;
    float dz = static_cast<float>(dzs);
    int nzs = static_cast<int>(ceil((zMax - zMin) / dz));
    headerFile << nx << " " << ny << " " << nzs << std::endl;
    headerFile << std::setprecision(10);
    headerFile << dx * horScale << " " << dy * horScale << " " << dz * vertScale << std::endl;
    double x0 = xmin + 0.5 * dx;
    double y0 = ymin + 0.5 * dy;
    double z0 = zMin + 0.5 * dz;
    headerFile << x0 * horScale << " " << y0 * horScale << " " << z0 * vertScale << std::endl;
    headerFile << angle << " 0\n";
    headerFile << kRMissing << std::endl;
    fName = std::string(fileName) + ".Sgri";
    headerFile << fName << std::endl;
    headerFile << "0\n";
    std::ofstream binFile;
    NRLib::OpenWrite(binFile, fName, std::ios::out | std::ios::binary);
    for (int k = 0; k < nzs; k++)
      for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
          double zTop = top[(size_t)j * nx + i], zBot = bot[(size_t)j * nx + i];
          double z = zMin + k * dz;
          float value;
          if (z < zTop || z > zBot) {
            value = static_cast<float>(kRMissing);
          } else {
            int simboxK = static_cast<int>((z - zTop) / (dzs * b.rel_thick(i, j)) + 0.5);
            value = b.value(i, j, simboxK);
          }
          NRLib::WriteBinaryFloat(binFile, value);
        }
    return 0;
  } catch (...) {
    return 1;
  }
}

// FFTGrid::writeStormFile, src/fftgrid.cpp:2118-2183, with the header of
// Simbox::getStormHeader, src/simbox.cpp:408-430, statement for statement on
// NRLib::ToString and NRLib::WriteBinaryFloat.  Writes path + ".storm"
// (binary) or path + ".txt" (ascii; fixed, 6 decimals, or scientific).
// grid: real padded float [nzp][nyp][rnxp]; padding writes the padded cube
// (nx, ny, nz are then nxp, nyp, nzp); the values are getRealValue(i, j, k,
// true) of the cube written.
int ref_write_storm(const char *path, const float *grid, int nx, int ny, int nz, int rnxp, int nyp, int cubetype,
                    double xmin, double lx, double ymin, double ly, double lz, double angle, const char *topName,
                    const char *botName, int flat, int ascii, int scientific) {
  try {
    const std::string top(topName), bot(botName);
    std::string header;
    if (ascii == 0)
      header = "storm_petro_binary\n";
    else
      header = "storm_petro_ascii\n";
    header += "0 " + NRLib::ToString(cubetype) + " " + NRLib::ToString(RMISSING_STORM, 6) + "\n";
    header += "FFTGrid\n";
    if (flat == 0)
      header += NRLib::ToString(xmin, 6) + " " + NRLib::ToString(lx, 6) + " " + NRLib::ToString(ymin, 6) + " " +
                NRLib::ToString(ly, 6) + " " + top + " " + bot + " 0.0 0.0\n";
    else
      header += NRLib::ToString(xmin, 6) + " " + NRLib::ToString(lx, 6) + " " + NRLib::ToString(ymin, 6) + " " +
                NRLib::ToString(ly, 6) + " 0.0 " + NRLib::ToString(lz, 6) + " 0.0 0.0\n";
    header += NRLib::ToString(lz, 6) + " " + NRLib::ToString(angle * 180 / NRLib::Pi, 6) + "\n\n";
    header += NRLib::ToString(nx) + " " + NRLib::ToString(ny) + " " + NRLib::ToString(nz) + "\n";
    const auto value = [&](int i, int j, int k) {
      return grid[(size_t)i + (size_t)rnxp * ((size_t)j + (size_t)nyp * k)];
    };
    if (ascii == 0) {
      std::ofstream binFile(std::string(path) + ".storm", std::ios::out | std::ios::binary);
      binFile << header;
      for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
          for (int i = 0; i < nx; i++) NRLib::WriteBinaryFloat(binFile, value(i, j, k));
      binFile << "0\n";
    } else {
      std::ofstream file(std::string(path) + ".txt");
      file << header;
      if (scientific)
        file << std::scientific;
      else
        file << std::fixed << std::setprecision(6);
      for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++) {
          for (int i = 0; i < nx - 1; i++) file << value(i, j, k) << " ";
          file << value(nx - 1, j, k) << "\n";
        }
      file << "0\n";
    }
    return 0;
  } catch (...) {
    return 1;
  }
}

}  // extern "C"
