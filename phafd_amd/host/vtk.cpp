// vtk.cpp -- the reference's dump file formats for one subdomain holding the
// whole box (Dump, src/dump.cpp; ReadDump, src/read_dump.cpp):
//   .vti  ImageData, one Float64 array per attribute in a raw appended block,
//         each preceded by its byte length as a 32-bit unsigned integer
//   .vtp  PolyData of the owned atoms, ASCII arrays
//   .pvti / .pvtp  the parallel headers naming the one piece
//   .pvd  the collection of timesteps
// Numbers are formatted as the reference's std::ostream prints them (the
// default format is printf's %.6g; "fixed" / "scientific" with a precision
// are %.Nf / %.Ne), here with std::to_chars, which is specified to give the
// same characters as those printf conversions.  Host only, no GPU work.
#include <algorithm>
#include <charconv>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "phafd_host.h"

namespace {

thread_local std::string g_err;

template <class F>
int guard(F &&f) {
  try {
    return f();
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// the stream format of the dump (src/dump.cpp:365-371): fixed / scientific
// with the given precision, scientific with 6 digits, or the default %.6g
struct NumFmt {
  std::chars_format f = std::chars_format::general;
  int prec = 6;
};
NumFmt num_fmt(int precision, int fmt) {
  NumFmt n;
  if (precision > 0 && fmt == PHAFD_VTK_FMT_LF) {
    n.f = std::chars_format::fixed;
    n.prec = precision;
  } else if (precision > 0 && fmt == PHAFD_VTK_FMT_E) {
    n.f = std::chars_format::scientific;
    n.prec = precision;
  } else if (fmt == PHAFD_VTK_FMT_E) {
    n.f = std::chars_format::scientific;
  }
  return n;
}

struct Out {
  std::string s;
  NumFmt nf;
  Out &operator<<(const char *t) { s += t; return *this; }
  Out &operator<<(const std::string &t) { s += t; return *this; }
  Out &operator<<(long long v) {
    char b[32];
    auto r = std::to_chars(b, b + sizeof b, v);
    s.append(b, r.ptr);
    return *this;
  }
  Out &operator<<(int v) { return *this << (long long)v; }
  Out &operator<<(double v) {
    char b[400];   // fixed format of the largest doubles
    auto r = std::to_chars(b, b + sizeof b, v, nf.f, nf.prec);
    if (r.ec != std::errc()) throw std::runtime_error("number formatting failed");
    s.append(b, r.ptr);
    return *this;
  }
};

struct File {
  FILE *f = nullptr;
  std::string path;
  File(const char *p, const char *mode) : path(p ? p : "") {
    if (!p) throw std::invalid_argument("null file name");
    f = std::fopen(p, mode);
    if (!f) throw std::runtime_error("Cannot open file " + path);
  }
  ~File() { if (f) std::fclose(f); }
  void put(const void *d, size_t n) {
    if (n && std::fwrite(d, 1, n, f) != n) throw std::runtime_error("write failed: " + path);
  }
  void put(const std::string &s) { put(s.data(), s.size()); }
  void close() {
    FILE *g = f;
    f = nullptr;
    if (std::fclose(g) != 0) throw std::runtime_error("write failed: " + path);
  }
};

enum Kind { VEC3, INT1 };
// the per-atom quantities of the atom dump (src/dump.cpp:985-1004); c_<id>:
// a per-atom compute of three components (compute pair), written as the
// reference writes a compute's array (src/dump.cpp:872-930, 1061-1133)
Kind atom_kind(const std::string &w) {
  if (w == "x" || w == "F" || w == "ux") return VEC3;
  if (w == "ix" || w == "ID" || w == "molID" || w == "type" || w == "label") return INT1;
  if (w.rfind("c_", 0) == 0 && w.size() > 2) return VEC3;
  if (w.rfind("f_", 0) == 0) throw std::invalid_argument("fix attributes are not dumped here: " + w);
  throw std::invalid_argument("Invalid per atom quantity in dump file.");
}

std::vector<std::string> name_list(int nattr, const char *const *names) {
  if (nattr <= 0 || !names) throw std::invalid_argument("Need attributes to write dump files.");
  std::vector<std::string> v;
  for (int i = 0; i < nattr; i++) {
    if (!names[i]) throw std::invalid_argument("null attribute name");
    v.emplace_back(names[i]);
  }
  return v;
}

const char *XML_HEAD = "<?xml version=\"1.0\"?>\n";

void grid_geometry_check(int nx, int ny, int nz) {
  if (nx <= 0 || ny <= 0 || nz <= 0) throw std::invalid_argument("bad grid size");
}

}  // namespace

extern "C" {

const char *phafd_host_last_error(void) { return g_err.c_str(); }

// Dump::write_grid_timestep (src/dump.cpp:681-780) with one rank: extent
// 0..Nz-1 and a byte length of Nx*Ny*Nz doubles, the size ReadDump expects
// of a one-rank file (src/read_dump.cpp:318-338)
int phafd_vtk_write_grid(const char *path, int nx, int ny, int nz, const double *origin, const double *spacing,
                         int nattr, const char *const *names, const double *const *arrays) {
  return guard([&] {
    grid_geometry_check(nx, ny, nz);
    auto at = name_list(nattr, names);
    if (!arrays) throw std::invalid_argument("null arrays");
    const uint64_t bl = (uint64_t)nx * ny * nz * sizeof(double);
    if (bl > UINT32_MAX) throw std::invalid_argument("grid too large for a 32-bit VTK byte length");
    const unsigned int bytelength = (unsigned int)bl;
    if ((uint64_t)(nattr - 1) * (bl + sizeof(bytelength)) > (uint64_t)INT_MAX)
      throw std::invalid_argument("appended offsets exceed the format's int range");
    const int zstart = 0, zend = nz - 1;
    Out o;
    o << XML_HEAD << "<VTKFile type=\"ImageData\"  version=\"1.0\" byte_order=\"LittleEndian\">\n"
      << "<ImageData WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << " " << zstart << "  " << zend << "\" "
      << "Origin=\"" << origin[0] << " " << origin[1] << " " << origin[2] << " \" "
      << "Spacing=\"" << spacing[0] << " " << spacing[1] << " " << spacing[2] << " \">\n"
      << "<Piece Extent=\"0 " << nx - 1 << " 0 " << ny - 1 << " " << zstart << " " << zend << "\">\n"
      << "<PointData Scalars=\"scalars\">\n";
    for (int k = 0; k < nattr; k++)
      o << "<DataArray Name=\"" << at[k] << "\" type=\"Float64\" format=\"appended\" offset=\""
        << (long long)((uint64_t)k * (bl + sizeof(bytelength))) << "\"/>\n";
    o << "</PointData>\n</Piece>\n</ImageData>\n<AppendedData encoding=\"raw\">\n_";
    File f(path, "wb");
    f.put(o.s);
    for (int k = 0; k < nattr; k++) {
      if (!arrays[k]) throw std::invalid_argument("null array for " + at[k]);
      f.put(&bytelength, sizeof(bytelength));
      f.put(arrays[k], bl);
    }
    f.put(std::string("\n</AppendedData>\n</VTKFile>\n"));
    f.close();
    return 0;
  });
}

// Dump::write_grid_timestep_pvti (src/dump.cpp:782-855), one piece
int phafd_vtk_write_pgrid(const char *path, int nx, int ny, int nz, const double *origin, const double *spacing,
                          int nattr, const char *const *names, const char *source) {
  return guard([&] {
    grid_geometry_check(nx, ny, nz);
    auto at = name_list(nattr, names);
    if (!source) throw std::invalid_argument("null piece source");
    Out o;
    o << XML_HEAD << "<VTKFile type=\"PImageData\"  version=\"1.0\" byte_order=\"LittleEndian\">\n"
      << "<PImageData WholeExtent=\"0 " << nx - 1 << " 0 " << ny - 1 << "  0 " << nz - 1 << "\" "
      << "Origin=\"" << origin[0] << " " << origin[1] << " " << origin[2] << " \" "
      << "Spacing=\"" << spacing[0] << " " << spacing[1] << " " << spacing[2] << "\" GhostLevel=\"1\">\n"
      << "<PPointData Scalars=\"scalars\">\n";
    for (auto &w : at) o << "<PDataArray Name=\"" << w << "\" type=\"Float64\"/>\n";
    o << "</PPointData>\n"
      << "<Piece Extent=\"0 " << nx << " 0 " << ny << " " << 0 << " " << nz - 1 << "\" Source=\"" << source
      << "\"/>\n"
      << "\n</PImageData>\n</VTKFile>\n";
    File f(path, "wb");
    f.put(o.s);
    f.close();
    return 0;
  });
}

// Dump::write_atom_timestep for atom/owned (src/dump.cpp:348-418, the ASCII
// arrays of src/dump.cpp:1014-1133): positions first, then the other
// attributes in the given order
int phafd_vtk_write_atoms(const char *path, int n, int nattr, const char *const *names, const void *const *arrays,
                          int precision, int fmt) {
  return guard([&] {
    auto at = name_list(nattr, names);
    if (n < 0 || !arrays) throw std::invalid_argument("bad atom arrays");
    if (fmt != PHAFD_VTK_FMT_LF && fmt != PHAFD_VTK_FMT_E)
      throw std::invalid_argument("Dump fmt must be either lf or e.");
    int ix = -1;
    for (int k = 0; k < nattr; k++) {
      atom_kind(at[k]);
      if (at[k] == "x" && ix < 0) ix = k;
      if (n > 0 && !arrays[k]) throw std::invalid_argument("null array for " + at[k]);
    }
    if (ix < 0) throw std::runtime_error("Atom dump must contain positions.");
    // the values of each array are formatted in atom chunks on several
    // threads, then written in order: the file is the sequential one
    const NumFmt nf = num_fmt(precision, fmt);
    const int nchunk = n < 8192 ? 1 : std::min(64, n / 4096);
    std::vector<int> order{ix};
    for (int k = 0; k < nattr; k++)
      if (at[k] != "x") order.push_back(k);
    const int ntask = (int)order.size() * nchunk;
    std::vector<std::string> body(ntask);
    auto task = [&](int t) {
      const int k = order[t / nchunk], c = t % nchunk;
      const int i0 = (int)((int64_t)n * c / nchunk), i1 = (int)((int64_t)n * (c + 1) / nchunk);
      Out o;
      o.nf = nf;
      if (atom_kind(at[k]) == VEC3) {
        o.s.reserve((size_t)(i1 - i0) * 3 * 16);
        const double *v = static_cast<const double *>(arrays[k]);
        for (int i = i0; i < i1; i++) o << v[3 * i] << " " << v[3 * i + 1] << " " << v[3 * i + 2] << " ";
      } else {
        o.s.reserve((size_t)(i1 - i0) * 8);
        const int *v = static_cast<const int *>(arrays[k]);
        for (int i = i0; i < i1; i++) o << v[i] << " ";
      }
      body[t].swap(o.s);
    };
    const int nthr = std::max(1, std::min({ntask, 16, (int)std::thread::hardware_concurrency()}));
    if (nthr == 1) {
      for (int t = 0; t < ntask; t++) task(t);
    } else {
      std::vector<std::thread> pool;
      std::vector<std::string> errs(nthr);
      for (int r = 0; r < nthr; r++)
        pool.emplace_back([&, r] {
          try {
            for (int t = r; t < ntask; t += nthr) task(t);
          } catch (const std::exception &e) {
            errs[r] = e.what();
          }
        });
      for (auto &th : pool) th.join();
      for (auto &e : errs)
        if (!e.empty()) throw std::runtime_error(e);
    }
    auto head = [&](int k) {
      Out o;
      const std::string &w = at[k];
      if (atom_kind(w) == VEC3)
        o << "<DataArray Name=\"" << w << "\" type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
      else
        o << "<DataArray Name=\"" << w << "\" type=\"Int64\" NumberOfComponents=\"1\" format=\"ascii\">\n";
      return o.s;
    };
    Out o;
    o << XML_HEAD << "<VTKFile type=\"PolyData\"  version=\"1.0\" byte_order=\"LittleEndian\">\n"
      << "<PolyData>\n<Piece NumberOfPoints=\"" << n << "\" NumberOfLines=\"1\">\n<Points>\n";
    File f(path, "wb");
    f.put(o.s);
    for (size_t j = 0; j < order.size(); j++) {
      f.put(head(order[j]));
      for (int c = 0; c < nchunk; c++) f.put(body[j * nchunk + c]);
      f.put(std::string("\n</DataArray>\n"));
      if (j == 0) f.put(std::string("</Points>\n<PointData>\n"));
    }
    f.put(std::string("</PointData>\n</Piece>\n</PolyData>\n</VTKFile>\n"));
    f.close();
    return 0;
  });
}

// Dump::write_atom_timestep_pvtp (src/dump.cpp:422-483), one piece
int phafd_vtk_write_patoms(const char *path, int nattr, const char *const *names, const char *source) {
  return guard([&] {
    auto at = name_list(nattr, names);
    if (!source) throw std::invalid_argument("null piece source");
    auto pdecl = [&](const std::string &w) {
      Out o;
      if (atom_kind(w) == VEC3)
        o << "<PDataArray Name=\"" << w << "\" type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\"/>\n";
      else
        o << "<PDataArray Name=\"" << w << "\" type=\"Int64\" NumberOfComponents=\"1\" format=\"ascii\"/>\n";
      return o.s;
    };
    Out o;
    o << XML_HEAD << "<VTKFile type=\"PPolyData\"  version=\"1.0\" byte_order=\"LittleEndian\">\n"
      << "<PPolyData GhostLevel=\"0\">\n<PPoints>\n" << pdecl("x") << "</PPoints>\n<PPointData>\n";
    for (auto &w : at)
      if (w != "x") o << pdecl(w);
    o << "</PPointData>\n<Piece Source=\"" << source << "\"/>\n</PPolyData>\n</VTKFile>\n";
    File f(path, "wb");
    f.put(o.s);
    f.close();
    return 0;
  });
}

// Dump::write_lammps_dump (src/dump.cpp:485-679): one timestep appended to
// a LAMMPS-style text file (the dump setup truncates it, src/dump.cpp:333-346);
// the box bounds are printed with the same stream flags as the values
int phafd_lammps_write_step(const char *path, int64_t timestep, int n, const double *boxlo, const double *boxhi,
                            int nattr, const char *const *names, const void *const *arrays, int precision, int fmt) {
  return guard([&] {
    auto at = name_list(nattr, names);
    if (n < 0 || !arrays || !boxlo || !boxhi) throw std::invalid_argument("bad arguments");
    if (fmt != PHAFD_VTK_FMT_LF && fmt != PHAFD_VTK_FMT_E)
      throw std::invalid_argument("Dump fmt must be either lf or e.");
    std::vector<Kind> kind;
    for (int k = 0; k < nattr; k++) {
      kind.push_back(atom_kind(at[k]));
      if (n > 0 && !arrays[k]) throw std::invalid_argument("null array for " + at[k]);
    }
    Out o;
    o.nf = num_fmt(precision, fmt);
    o.s.reserve((size_t)n * 16 * (3 * nattr) + 1024);
    o << "ITEM: TIMESTEP\n" << (long long)timestep << "\n" << "ITEM: NUMBER OF ATOMS\n" << n << "\n"
      << "ITEM: BOX BOUNDS pp pp pp\n";
    for (int d = 0; d < 3; d++) o << boxlo[d] << " " << boxhi[d] << "\n";
    o << "ITEM: ATOMS ";
    for (int k = 0; k < nattr; k++) {
      const std::string &w = at[k];
      if (w == "x") o << "x y z ";
      else if (w == "F") o << "Fx Fy Fz ";
      else if (w == "ux") o << "ux uy uz ";
      else if (kind[k] == VEC3) o << w << "_0 " << w << "_1 " << w << "_2 ";   // a compute of 3 components
      else o << w << " ";
    }
    o << "\n";
    for (int i = 0; i < n; i++) {
      for (int k = 0; k < nattr; k++) {
        if (kind[k] == VEC3) {
          const double *v = static_cast<const double *>(arrays[k]) + 3 * (size_t)i;
          o << v[0] << " " << v[1] << " " << v[2] << " ";
        } else {
          o << static_cast<const int *>(arrays[k])[i] << " ";
        }
      }
      o << "\n";
    }
    File f(path, "ab");
    f.put(o.s);
    f.close();
    return 0;
  });
}

// Dump::write_collection_header / _middle / _footer (src/dump.cpp:235-313)
int phafd_vtk_collection(const char *path, int part, double time, const char *file) {
  return guard([&] {
    Out o;
    const char *mode = "ab";
    if (part == PHAFD_VTK_PVD_HEADER) {
      o << XML_HEAD << "<VTKFile type=\"Collection\"  version=\"1.0\" byte_order=\"LittleEndian\">\n<Collection>\n";
      mode = "wb";
    } else if (part == PHAFD_VTK_PVD_DATASET) {
      if (!file) throw std::invalid_argument("null data set file");
      o << "<DataSet timestep=\"" << time << "\" group=\"\" part=\"0\" file=\"" << file << "\"/>\n";
    } else if (part == PHAFD_VTK_PVD_FOOTER) {
      o << "</Collection>\n</VTKFile>";
    } else {
      throw std::invalid_argument("unknown collection part");
    }
    File f(path, mode);
    f.put(o.s);
    f.close();
    return 0;
  });
}

// ReadDump::process_grid_attributes + read_binary_data (src/read_dump.cpp:
// 103-195, 300-338) for one rank (no neighbour planes, with or without
// no_padding): out[k] receives requested attribute k
int phafd_vtk_read_grid(const char *path, int nx, int ny, int nz, int nattr, const char *const *names,
                        double *const *out, int no_padding) {
  (void)no_padding;   // one rank: no padding planes either way
  return guard([&] {
    grid_geometry_check(nx, ny, nz);
    if (nattr <= 0 || !names) throw std::invalid_argument("Need attributes to read dump files.");
    auto at = name_list(nattr, names);
    for (auto &w : at)
      if (w.rfind("c_", 0) == 0 || w.rfind("f_", 0) == 0) throw std::runtime_error("Cannot read computes from dump");
    bool has_phi = false;
    for (auto &w : at) has_phi |= (w == "phi");
    if (!has_phi) throw std::runtime_error("Read dump needs attribute phi.");
    std::ifstream f(path ? path : "", std::ios::in | std::ios::binary);
    if (!path || f.fail()) throw std::runtime_error("file not found.");
    const std::string expstart = "<DataArray Name=";
    std::vector<int> found;   // index into at, or -1 for an attribute not asked for
    int ignored = 0;
    std::string line;
    while (std::getline(f, line)) {
      if (line == "<AppendedData encoding=\"raw\">") break;
      bool hit = false;
      for (int k = 0; k < nattr; k++) {
        const std::string e = "<DataArray Name=\"" + at[k] + "\" type=\"Float64\" format=\"appended\" offset=\"";
        if (line.compare(0, e.size(), e) == 0) {
          found.push_back(k);
          hit = true;
          break;
        }
      }
      if (!hit && line.compare(0, expstart.size(), expstart) == 0) {
        found.push_back(-1);
        ignored++;
      }
    }
    if ((int)found.size() - ignored < nattr) {
      std::string msg = "Not all attributes requested in read_dump command are in " + std::string(path) +
                        ". List of attributes in file are:";
      for (auto &w : at) msg += " " + w;
      throw std::runtime_error(msg + ".");
    }
    char us;
    f.read(&us, 1);
    const uint64_t want = (uint64_t)nx * ny * nz * sizeof(double);
    for (int k : found) {
      unsigned int bl = 0;
      f.read(reinterpret_cast<char *>(&bl), sizeof(bl));
      if (!f) throw std::runtime_error("truncated dump file " + std::string(path));
      if (bl != want) {
        if (k < 0) throw std::runtime_error("incorrect array size in " + std::string(path));
        throw std::runtime_error("incorrect size " + at[k] + " in file.");
      }
      if (k < 0 || !out || !out[k]) {
        f.ignore((std::streamsize)bl);
      } else {
        f.read(reinterpret_cast<char *>(out[k]), (std::streamsize)bl);
      }
      if (!f) throw std::runtime_error("truncated dump file " + std::string(path));
    }
    return 0;
  });
}

// ReadDump::process_atom_attributes + read_ascii_data (src/read_dump.cpp:
// 199-296): the line after each matching array header holds natoms values
// (three per atom for x, F, ux); returns natoms, the owned count
int phafd_vtk_read_atoms(const char *path, int natoms, int nattr, const char *const *names, int molecule,
                         void *const *out) {
  return guard([&] {
    if (natoms < 0) throw std::invalid_argument("bad atom count");
    if (nattr <= 0 || !names) throw std::invalid_argument("Need attributes to read dump files.");
    auto at = name_list(nattr, names);
    for (auto &w : at)
      if (w.rfind("c_", 0) == 0 || w.rfind("f_", 0) == 0) throw std::runtime_error("Cannot read computes from dump");
    std::vector<std::string> need = {"x", "type", "ID"};
    if (molecule) { need.push_back("molID"); need.push_back("ix"); }
    for (auto &w : need) {
      bool has = false;
      for (auto &a : at) has |= (a == w);
      if (!has) throw std::runtime_error("Read dump needs attribute " + w + ".");
    }
    std::ifstream f(path ? path : "");
    if (!path || f.fail()) throw std::runtime_error("file not found.");
    std::vector<std::string> found;
    std::string line;
    while (std::getline(f, line)) {
      for (int k = 0; k < nattr; k++) {
        const std::string &w = at[k];
        const bool vec = line == "<DataArray Name=\"" + w + "\" type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">";
        const bool scal = line == "<DataArray Name=\"" + w + "\" type=\"Int64\" NumberOfComponents=\"1\" format=\"ascii\">";
        if (!vec && !scal) continue;
        std::getline(f, line);
        const Kind kind = atom_kind(w);
        if ((vec && kind != VEC3) || (scal && kind != INT1))
          throw std::runtime_error("Attribute " + w + " not recognised.");
        const char *p = line.c_str();
        char *end = nullptr;
        if (kind == VEC3) {
          double *v = out ? static_cast<double *>(out[k]) : nullptr;
          for (int i = 0; i < 3 * natoms; i++) {
            double d = std::strtod(p, &end);
            if (end == p) throw std::runtime_error("short array " + w + " in " + std::string(path));
            if (v) v[i] = d;
            p = end;
          }
        } else {
          int *v = out ? static_cast<int *>(out[k]) : nullptr;
          for (int i = 0; i < natoms; i++) {
            long d = std::strtol(p, &end, 10);
            if (end == p) throw std::runtime_error("short array " + w + " in " + std::string(path));
            if (v) v[i] = (int)d;
            p = end;
          }
        }
        found.push_back(w);
        break;
      }
    }
    if ((int)found.size() != nattr) {
      std::string msg = "Only found attributes ";
      for (auto &w : found) msg += w + ", ";
      throw std::runtime_error(msg);
    }
    return natoms;
  });
}

}  // extern "C"
