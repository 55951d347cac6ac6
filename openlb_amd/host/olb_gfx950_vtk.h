/*
 * olb_gfx950_vtk.h -- SuperVTMwriter3D of the C++ host mirror: VTK output of
 * one lattice (one cuboid) in the layout of io/superVtmWriter3D.hh, so the
 * files open in ParaView exactly as the reference's do.
 *
 *   DIR/NAME.pvd                       master collection (createMasterFile,
 *                                      dataPVDmaster: closing 25 bytes rewritten)
 *   DIR/data/NAME_iT%07d.vtm           vtkMultiBlockDataSet of the pieces
 *   DIR/data/NAME_iT%07diC%05d.vti     ImageData, WholeExtent -overlap ..
 *                                      extent+overlap-1, one Float32 DataArray
 *                                      per functor, x fastest, components
 *                                      interleaved; zlib + base64 behind the
 *                                      prefix {1, raw, raw, compressed}
 *                                      (compress, the default), base64 behind
 *                                      the byte count (binary), or text
 *
 * Functors: velocity and density (SuperLatticeVelocity3D / Density3D) and
 * physVelocity / physPressure (SuperLatticePhysVelocity3D / PhysPressure3D)
 * through a converter's velocity and pressure factors.  Values come from
 * SuperLattice::computeRhoU (the momenta of each cell's dynamics); overlap
 * points on a periodic axis take their periodic image's value, as the
 * reference's communicate() before writing makes them.  Needs zlib (-lz).
 */
#ifndef OLB_GFX950_VTK_H
#define OLB_GFX950_VTK_H

#include <zlib.h>

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "olb_gfx950.h"

namespace olb {

/// the converter factors the physical functors use (core/unitConverter.h)
template <typename T>
struct VtkConversion {
  T velocity = 1;          // getConversionFactorVelocity
  T pressure = 1;          // getConversionFactorPressure
  T charPhysPressure = 0;  // getPhysPressure(p) = pressure * p + charPhysPressure
};

enum class VtkFunctor { Velocity, Density, PhysVelocity, PhysPressure };

template <typename T>
class SuperVTMwriter3D {
public:
  SuperVTMwriter3D(std::string name, std::string vtkOutDir = "./tmp/vtkData/", int overlap = 1,
                   bool binary = true, bool compress = true)
    : _name(std::move(name)), _dir(std::move(vtkOutDir)), _overlap(overlap), _binary(binary), _compress(compress)
  {
    if (!_dir.empty() && _dir.back() != '/') _dir += '/';
  }

  void setConversion(const VtkConversion<T>& c) { _conv = c; }
  void addFunctor(VtkFunctor f) { _functors.push_back(f); }

  void createMasterFile()
  {
    makeDirs();
    std::ofstream out(_dir + _name + ".pvd", std::ios::trunc);
    out << "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
        << "<Collection>\n" << kClosePvd;
  }

  template <typename DESCRIPTOR>
  void write(SuperLattice<T, DESCRIPTOR>& lattice, int iT)
  {
    if (_functors.empty()) throw std::runtime_error("No functor to write");
    makeDirs();
    const std::string vtm = fileName(iT, -1) + ".vtm";
    {
      std::fstream pvd(_dir + _name + ".pvd", std::ios::in | std::ios::out | std::ios::ate);
      if (pvd) {
        pvd.seekp(-(std::streamoff)std::string(kClosePvd).size(), std::ios::end);
        pvd << "<DataSet timestep=\"" << iT << "\" group=\"\" part=\"\" file=\"data/" << vtm << "\"/>\n" << kClosePvd;
      }
    }
    {
      std::ofstream out(_dir + "data/" + vtm, std::ios::trunc);
      out << "<?xml version=\"1.0\"?>\n<VTKFile type=\"vtkMultiBlockDataSet\" version=\"1.0\" "
          << "byte_order=\"LittleEndian\">\n<vtkMultiBlockDataSet>\n"
          << "<Block index=\"0\" >\n<DataSet index= \"0\" file=\"" << fileName(iT, 0) << ".vti\">\n</DataSet>\n</Block>\n"
          << "</vtkMultiBlockDataSet>\n</VTKFile>\n";
    }
    SuperGeometry<T>& g = lattice.getGeometry();
    std::vector<T> rho, u;
    lattice.computeRhoU(rho, u);
    const auto e = g.getExtent();
    const auto per = g.getPeriodicity();
    if (_overlap > g.getOverlap()) throw std::invalid_argument("the overlap written exceeds the lattice padding");
    std::ofstream out(_dir + "data/" + fileName(iT, 0) + ".vti", std::ios::trunc);
    const auto origin = g.getPhysR(0, 0, 0);
    std::ostringstream ext;
    ext << -_overlap << " " << e[0] + _overlap - 1 << " " << -_overlap << " " << e[1] + _overlap - 1 << " " << -_overlap
        << " " << e[2] + _overlap - 1;
    out << "<?xml version=\"1.0\"?>\n<VTKFile type=\"ImageData\" version=\"0.1\" ";
    out << (_compress ? "byte_order=\"LittleEndian\" compressor=\"vtkZLibDataCompressor\">\n"
                      : "byte_order=\"LittleEndian\">\n");
    out << "<ImageData WholeExtent=\"" << ext.str() << "\" Origin=\"" << origin[0] << " " << origin[1] << " "
        << origin[2] << "\" Spacing=\"" << g.getDeltaR() << " " << g.getDeltaR() << " " << g.getDeltaR() << "\">\n";
    out << "<Piece Extent=\"" << ext.str() << "\">\n<PointData>\n";
    for (VtkFunctor f : _functors) {
      const int dim = (f == VtkFunctor::Velocity || f == VtkFunctor::PhysVelocity) ? 3 : 1;
      std::vector<float> v;
      v.reserve((std::size_t)dim * (e[0] + 2 * _overlap) * (e[1] + 2 * _overlap) * (e[2] + 2 * _overlap));
      for (int z = -_overlap; z < e[2] + _overlap; ++z) {
        for (int y = -_overlap; y < e[1] + _overlap; ++y) {
          for (int x = -_overlap; x < e[0] + _overlap; ++x) {
            const int r[3] = {x, y, z};
            int w[3];
            for (int a = 0; a < 3; ++a) w[a] = (per[a] && (r[a] < 0 || r[a] >= e[a])) ? (r[a] + e[a]) % e[a] : r[a];
            const std::size_t c = g.getCellId(w[0], w[1], w[2]);
            switch (f) {
            case VtkFunctor::Velocity:
              for (int a = 0; a < 3; ++a) v.push_back(float(u[3 * c + a]));
              break;
            case VtkFunctor::Density:
              v.push_back(float(rho[c]));
              break;
            case VtkFunctor::PhysVelocity:
              for (int a = 0; a < 3; ++a) v.push_back(float(_conv.velocity * u[3 * c + a]));
              break;
            case VtkFunctor::PhysPressure:
              // (computeRho() - 1.0) / invCs2 in double, held in T (latticePhysPressure3D.hh:73-76)
              v.push_back(float(_conv.pressure * T(((double)rho[c] - 1.0) / (double)T(3)) + _conv.charPhysPressure));
              break;
            }
          }
        }
      }
      dataArray(out, name(f), dim, v);
    }
    out << "</PointData>\n</Piece>\n</ImageData>\n</VTKFile>\n";
  }

private:
  static constexpr const char* kClosePvd = "</Collection>\n</VTKFile>\n";

  static const char* name(VtkFunctor f)
  {
    switch (f) {
    case VtkFunctor::Velocity: return "velocity";
    case VtkFunctor::Density: return "density";
    case VtkFunctor::PhysVelocity: return "physVelocity";
    default: return "physPressure";
    }
  }

  void makeDirs() const
  {
    ::mkdir(_dir.c_str(), 0755);
    ::mkdir((_dir + "data/").c_str(), 0755);
  }

  std::string fileName(int iT, int iC) const
  {
    char buf[64];
    if (iC < 0) {
      std::snprintf(buf, sizeof(buf), "_iT%07d", iT);
    } else {
      std::snprintf(buf, sizeof(buf), "_iT%07diC%05d", iT, iC);
    }
    return _name + buf;
  }

  void dataArray(std::ofstream& out, const char* fname, int dim, const std::vector<float>& v) const
  {
    out << "<DataArray type=\"Float32\" Name=\"" << fname << "\" NumberOfComponents=\"" << dim << "\" ";
    out << ((_compress || _binary) ? "format=\"binary\" encoding=\"base64\">\n" : "format=\"ascii\" >\n");
    const auto* bytes = reinterpret_cast<const std::uint8_t*>(v.data());
    const std::uint32_t raw = (std::uint32_t)(v.size() * sizeof(float));
    if (_compress) {
      uLongf packed = compressBound(raw);
      std::vector<std::uint8_t> z(packed);
      if (compress2(z.data(), &packed, bytes, raw, -1) != Z_OK) throw std::runtime_error("zlib compress2 failed");
      const std::uint32_t prefix[4] = {1, raw, raw, (std::uint32_t)packed};
      out << base64(reinterpret_cast<const std::uint8_t*>(prefix), sizeof(prefix)) << base64(z.data(), packed);
    } else if (_binary) {
      out << base64(reinterpret_cast<const std::uint8_t*>(&raw), 4) << base64(bytes, raw);
    } else {
      for (float f : v) out << f << " ";
    }
    out << "\n</DataArray>\n";
  }

  std::string _name, _dir;
  int _overlap;
  bool _binary, _compress;
  VtkConversion<T> _conv;
  std::vector<VtkFunctor> _functors;
};

}  // namespace olb

#endif  // OLB_GFX950_VTK_H
