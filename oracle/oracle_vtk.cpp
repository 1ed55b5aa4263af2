// oracle_vtk.cpp -- TEST INFRASTRUCTURE ONLY: the reference's dump writers
// restated with std::ostream, as the reference formats them (Dump,
// src/dump.cpp), for one rank holding the whole box.  The product's writer
// (phafd_amd/host/vtk.cpp) formats numbers another way (std::to_chars); the
// tests compare the two files byte for byte.  Nothing in the product links
// or calls this file.
//
// One-rank grid files: the byte length written before each array is the
// size ReadDump requires of a one-rank file (Nx*Ny*Nz doubles,
// src/read_dump.cpp:318-338); the reference's writer sizes it for a
// neighbour plane (src/dump.cpp:730-733) that a single rank does not have,
// and its plane exchange addresses rank 1 (src/dump.cpp:1155-1158).
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <stdexcept>
#include <string>
#include <vector>

#include "oracle.h"

namespace {
thread_local std::string g_verr;

template <class F>
int vguard(F &&f) {
  try {
    return f();
  } catch (const std::exception &e) {
    g_verr = e.what();
    return -1;
  }
}

// x, F, ux (Matrix3Xd) and a per-atom compute of three components: the
// same header text and value layout (src/dump.cpp:1014-1133)
bool is_vec3(const std::string &w) { return w == "x" || w == "F" || w == "ux" || w.rfind("c_", 0) == 0; }
}  // namespace

extern "C" {

const char *orc_vtk_last_error(void) { return g_verr.c_str(); }

// Dump::write_grid_timestep, src/dump.cpp:681-780
int orc_vtk_write_grid(const char *path, int nx, int ny, int nz, const double *origin, const double *spacing,
                       int nattr, const char *const *names, const double *const *arrays) {
  return vguard([&] {
    std::ofstream f(path, std::ios::out | std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Cannot open file");
    const int zstart = 0, zend = nz - 1;
    unsigned int bytelength = (unsigned int)((uint64_t)nx * ny * nz * sizeof(double));
    f << "<?xml version=\"1.0\"?>" << std::endl
      << "<VTKFile type=\"ImageData\"  version=\"1.0\"" << " byte_order=\"LittleEndian\">" << std::endl
      << "<ImageData WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << " " << zstart << "  " << zend << "\" "
      << "Origin=\"" << origin[0] << " " << origin[1] << " " << origin[2] << " \" "
      << "Spacing=\"" << spacing[0] << " " << spacing[1] << " " << spacing[2] << " " << "\">" << std::endl
      << "<Piece Extent=\"0 " << nx - 1 << " 0 " << ny - 1 << " " << zstart << " " << zend << "\">" << std::endl
      << "<PointData Scalars=\"scalars\">" << std::endl;
    for (int k = 0; k < nattr; k++) {
      int offset = (int)(k * (bytelength + sizeof(bytelength)));
      f << "<DataArray Name=\"" << names[k] << "\" type=\"Float64\" format=\"appended\" " << "offset=\"" << offset
        << "\"/>" << std::endl;
    }
    f << "</PointData>" << std::endl
      << "</Piece>" << std::endl
      << "</ImageData>" << std::endl
      << "<AppendedData encoding=\"raw\">" << std::endl
      << "_";
    for (int k = 0; k < nattr; k++) {
      f.write((const char *)&bytelength, sizeof(bytelength));
      // row by row, as the reference writes its padded arrays
      for (int i = 0; i < nz; i++)
        for (int j = 0; j < ny; j++)
          f.write((const char *)(arrays[k] + ((size_t)i * ny + j) * nx), sizeof(double) * nx);
    }
    f << std::endl << "</AppendedData>" << std::endl << "</VTKFile>" << std::endl;
    return 0;
  });
}

// Dump::write_grid_timestep_pvti, src/dump.cpp:782-855
int orc_vtk_write_pgrid(const char *path, int nx, int ny, int nz, const double *origin, const double *spacing,
                        int nattr, const char *const *names, const char *source) {
  return vguard([&] {
    std::ofstream f(path, std::ios::out | std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Cannot open file");
    f << "<?xml version=\"1.0\"?>" << std::endl
      << "<VTKFile type=\"PImageData\"  version=\"1.0\"" << " byte_order=\"LittleEndian\">" << std::endl
      << "<PImageData WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << " " << " 0 " << nz - 1 << "\" "
      << "Origin=\"" << origin[0] << " " << origin[1] << " " << origin[2] << " \" "
      << "Spacing=\"" << spacing[0] << " " << spacing[1] << " " << spacing[2] << "\"" << " GhostLevel=\"1\">"
      << std::endl
      << "<PPointData Scalars=\"scalars\">" << std::endl;
    for (int k = 0; k < nattr; k++) f << "<PDataArray Name=\"" << names[k] << "\" type=\"Float64\"/>" << std::endl;
    f << "</PPointData>" << std::endl;
    const int zs = 0, ze = nz - 1;
    f << "<Piece Extent=\"0 " << nx << " 0 " << ny << " " << zs << " " << ze << "\" Source=\"" << source << "\"/>"
      << std::endl;
    f << std::endl << "</PImageData>" << std::endl << "</VTKFile>" << std::endl;
    return 0;
  });
}

// Dump::write_atom_timestep (atom/owned) with write_ascii_data,
// src/dump.cpp:348-418, 1014-1133
int orc_vtk_write_atoms(const char *path, int n, int nattr, const char *const *names, const void *const *arrays,
                        int precision, int fmt) {
  return vguard([&] {
    std::ofstream f(path, std::ios::out);
    if (!f.is_open()) throw std::runtime_error("Cannot open file");
    if (precision > 0 && fmt == 0) f << std::fixed << std::setprecision(precision);
    else if (precision > 0 && fmt == 1) f << std::scientific << std::setprecision(precision);
    else if (fmt == 1) f << std::scientific;
    auto one = [&](int k) {
      std::string w = names[k];
      if (is_vec3(w)) {
        f << "<DataArray Name=\"" << w << "\" " << "type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">"
          << std::endl;
        const double *v = (const double *)arrays[k];
        for (int i = 0; i < n; i++) f << v[3 * i] << " " << v[3 * i + 1] << " " << v[3 * i + 2] << " ";
      } else {
        f << "<DataArray Name=\"" << w << "\" " << "type=\"" << "Int64" << "\" NumberOfComponents=\"" << 1
          << "\" format=\"ascii\">" << std::endl;
        const int *v = (const int *)arrays[k];
        for (int i = 0; i < n; i++) f << v[i] << " ";
      }
      f << std::endl << "</DataArray>" << std::endl;
    };
    f << "<?xml version=\"1.0\"?>" << std::endl
      << "<VTKFile type=\"PolyData\"  version=\"1.0\"" << " byte_order=\"LittleEndian\">" << std::endl
      << "<PolyData>" << std::endl
      << "<Piece NumberOfPoints=\"" << n << "\" NumberOfLines=\"1\">" << std::endl;
    int ix = -1;
    for (int k = 0; k < nattr && ix < 0; k++)
      if (std::string(names[k]) == "x") ix = k;
    if (ix < 0) throw std::runtime_error("Atom dump must contain positions.");
    f << "<Points>" << std::endl;
    one(ix);
    f << "</Points>" << std::endl;
    f << "<PointData>" << std::endl;
    for (int k = 0; k < nattr; k++)
      if (std::string(names[k]) != "x") one(k);
    f << "</PointData>" << std::endl;
    f << "</Piece>" << std::endl << "</PolyData>" << std::endl << "</VTKFile>" << std::endl;
    return 0;
  });
}

// Dump::write_atom_timestep_pvtp, src/dump.cpp:422-483
int orc_vtk_write_patoms(const char *path, int nattr, const char *const *names, const char *source) {
  return vguard([&] {
    std::ofstream f(path, std::ios::out);
    if (!f.is_open()) throw std::runtime_error("Cannot open file");
    auto pdecl = [&](const std::string &w) {
      if (is_vec3(w))
        f << "<PDataArray Name=\"" << w << "\" " << "type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\"/>"
          << std::endl;
      else
        f << "<PDataArray Name=\"" << w << "\" " << "type=\"" << "Int64" << "\" NumberOfComponents=\"" << 1
          << "\" format=\"ascii\"/>" << std::endl;
    };
    f << "<?xml version=\"1.0\"?>" << std::endl
      << "<VTKFile type=\"PPolyData\"  version=\"1.0\"" << " byte_order=\"LittleEndian\">" << std::endl
      << "<PPolyData GhostLevel=\"0\">" << std::endl
      << "<PPoints>" << std::endl;
    pdecl("x");
    f << "</PPoints>" << std::endl;
    f << "<PPointData>" << std::endl;
    for (int k = 0; k < nattr; k++)
      if (std::string(names[k]) != "x") pdecl(names[k]);
    f << "</PPointData>" << std::endl;
    f << "<Piece Source=\"" + std::string(source) + std::string("\"/>") << std::endl;
    f << "</PPolyData>" << std::endl << "</VTKFile>" << std::endl;
    return 0;
  });
}

// Dump::write_lammps_dump, src/dump.cpp:485-679 (owned atoms, appended)
int orc_lammps_write_step(const char *path, int64_t timestep, int n, const double *boxlo, const double *boxhi,
                          int nattr, const char *const *names, const void *const *arrays, int precision, int fmt) {
  return vguard([&] {
    std::fstream f(path, std::ios::out | std::ios::app);
    if (!f.is_open()) throw std::runtime_error("Cannot open file");
    if (precision > 0 && fmt == 0) f << std::fixed << std::setprecision(precision);
    else if (precision > 0 && fmt == 1) f << std::scientific << std::setprecision(precision);
    else if (fmt == 1) f << std::scientific;
    f << "ITEM: TIMESTEP" << std::endl << timestep << std::endl;
    f << "ITEM: NUMBER OF ATOMS" << std::endl << n << std::endl;
    f << "ITEM: BOX BOUNDS pp pp pp" << std::endl
      << boxlo[0] << " " << boxhi[0] << std::endl
      << boxlo[1] << " " << boxhi[1] << std::endl
      << boxlo[2] << " " << boxhi[2] << std::endl;
    f << "ITEM: ATOMS ";
    for (int k = 0; k < nattr; k++) {
      std::string w = names[k];
      if (w == "x") f << "x y z ";
      else if (w == "F") f << "Fx Fy Fz ";
      else if (w == "ux") f << "ux uy uz ";
      else if (is_vec3(w))
        for (int c = 0; c < 3; c++) f << w << "_" << c << " ";
      else f << w << " ";
    }
    f << std::endl;
    for (int i = 0; i < n; i++) {
      for (int k = 0; k < nattr; k++) {
        if (is_vec3(names[k])) {
          const double *v = (const double *)arrays[k] + 3 * (size_t)i;
          f << v[0] << " " << v[1] << " " << v[2] << " ";
        } else {
          f << ((const int *)arrays[k])[i] << " ";
        }
      }
      f << std::endl;
    }
    return 0;
  });
}

// Dump::write_collection_header / _middle / _footer, src/dump.cpp:235-313
int orc_vtk_collection(const char *path, int part, double time, const char *file) {
  return vguard([&] {
    if (part == 0) {
      std::ofstream f(path);
      if (!f.is_open()) throw std::runtime_error("Cannot open file");
      f << "<?xml version=\"1.0\"?>" << std::endl
        << "<VTKFile type=\"Collection\"  version=\"1.0\"" << " byte_order=\"LittleEndian\">" << std::endl
        << "<Collection>" << std::endl;
    } else {
      std::fstream f(path, std::ios_base::app);
      if (!f.is_open()) throw std::runtime_error("Cannot open file");
      if (part == 1)
        f << "<DataSet timestep=\"" << time << "\" group=\"\" part=\"0\"" << " file=\"" << file << "\"/>"
          << std::endl;
      else
        f << "</Collection>" << std::endl << "</VTKFile>";
    }
    return 0;
  });
}

}  // extern "C"
