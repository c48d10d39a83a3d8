// smg_io.cpp -- the reference's output files for particles and meshes, single rank.
//
//  VTK XML unstructured grids with raw appended data (each array: a uint32 byte count,
//  then the values), a .pvtu naming the rank pieces and a .pvd collection that gains
//  one DataSet per output:
//    particles  Particle_binary_writer::Output_particle   particle_binary_writer.cpp:36-401
//               (point data id, local_id, type, h, NN, color, norm; one vertex cell per
//               particle)
//    meshes     Mesh_binary_writer::Output_mesh           mesh_binary_writer.cpp:36-650
//               (point data type, h, color; triangles, or tetrahedra with cell data
//               radius_ratio and mindihedangle)
//  Post files (SPH_mesh_generation::Save_particle_post_file, sph_mesh_generation.cpp:
//  2624-2667, read back by Load_data_for_post, :426-590): list.csv holds the number of
//  ranks, then one "step count" line per saved step; step_<step>_rank_<r>.pst holds an
//  int32 count and 3 doubles per particle.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "smg.h"
#include "smg_internal.hpp"

namespace smg {
namespace {

std::string join(const char* dir, const std::string& name) {
  std::string d = dir && dir[0] ? dir : ".";
  if (d.back() != '/') d += '/';
  return d + name;
}

// a DataArray header line and its appended block (byte count + values)
class Appended {
 public:
  template <typename T>
  void add(std::ostream& xml, const char* vtk_type, const char* name, int comps, const std::vector<T>& v,
           bool with_comps = true) {
    xml << "      <DataArray type=\"" << vtk_type << "\" Name=\"" << name << "\"";
    if (with_comps) xml << " NumberOfComponents=\"" << comps << "\"";
    xml << " format=\"appended\" offset=\"" << offset_ << "\"/>\n";
    const uint32_t bytes = uint32_t(v.size() * sizeof(T));
    data_.write(reinterpret_cast<const char*>(&bytes), sizeof(bytes));
    if (bytes) data_.write(reinterpret_cast<const char*>(v.data()), bytes);
    offset_ += int64_t(bytes) + int64_t(sizeof(bytes));
  }
  std::string str() const { return data_.str(); }

 private:
  std::ostringstream data_;
  int64_t offset_ = 0;
};

bool too_big(int64_t n, int per) { return n * per * 8 > int64_t(0xffffffffu); }

template <typename T>
std::vector<T> or_zero(const T* p, int64_t n, int per) {
  std::vector<T> v(size_t(n) * per, T(0));
  if (p) std::memcpy(v.data(), p, sizeof(T) * v.size());
  return v;
}

std::vector<float> to_float(const double* p, int64_t n, int per) {
  std::vector<float> v(size_t(n) * per, 0.f);
  if (p)
    for (size_t q = 0; q < v.size(); q++) v[q] = float(p[q]);
  return v;
}

int write_file(const std::string& path, const std::string& text) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) {
    set_error("cannot write " + path);
    return SMG_ERR_ARG;
  }
  f << text;
  f.close();
  if (!f) {
    set_error("write failed: " + path);
    return SMG_ERR_ARG;
  }
  return SMG_OK;
}

// the collection file: created at n_out == 0, otherwise the new DataSet replaces the
// two closing lines
int update_pvd(const std::string& path, int n_out, double time, const std::string& piece) {
  std::ostringstream ds;
  ds << "    <DataSet timestep=\"" << time << "\" group=\"\" part=\"0\" file=\"./" << piece << "\"/>\n";
  std::vector<std::string> lines;
  if (n_out > 0) {
    std::ifstream in(path);
    if (!in) {
      set_error("cannot read " + path + " (write output 0 first)");
      return SMG_ERR_STATE;
    }
    std::string line;
    while (std::getline(in, line)) lines.push_back(line);
    if (lines.size() < 4 || lines[lines.size() - 1].find("</VTKFile>") == std::string::npos) {
      set_error("not a collection file: " + path);
      return SMG_ERR_ARG;
    }
    lines.resize(lines.size() - 2);
  } else {
    lines = {"<VTKFile type=\"Collection\" version=\"0.1\">", "  <Collection>"};
  }
  std::ostringstream out;
  for (const auto& l : lines) out << l << "\n";
  out << ds.str() << "  </Collection>\n</VTKFile>\n";
  return write_file(path, out.str());
}

struct PointArrays {
  const char* type;
  const char* name;
  int comps;
};

int write_pvtu(const std::string& path, const std::vector<PointArrays>& point, const std::vector<PointArrays>& cell,
               const std::string& piece) {
  std::ostringstream x;
  x << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
    << " <PUnstructuredGrid GhostLevel=\"0\">\n  <PPointData Vectors=\"vector\">\n";
  for (const auto& a : point)
    x << "    <PDataArray type=\"" << a.type << "\" Name=\"" << a.name << "\" NumberOfComponents=\"" << a.comps
      << "\"/>\n";
  x << "  </PPointData>\n";
  if (!cell.empty()) {
    x << "  <PCellData Vectors=\"vector\">\n";
    for (const auto& a : cell)
      x << "    <PDataArray type=\"" << a.type << "\" Name=\"" << a.name << "\" NumberOfComponents=\"" << a.comps
        << "\"/>\n";
    x << "  </PCellData>\n";
  }
  x << "  <PPoints>\n    <PDataArray type=\"Float32\" Name=\"Points\" NumberOfComponents=\"3\"/>\n  </PPoints>\n"
    << "    <Piece Source=\"./" << piece << "\"/>\n </PUnstructuredGrid>\n</VTKFile>\n";
  return write_file(path, x.str());
}

int check_points(const smg_vtu_points* p) {
  if (!p || p->n < 0 || (p->n > 0 && !p->x)) {
    set_error("vtu writer: bad points");
    return SMG_ERR_ARG;
  }
  if (too_big(p->n, 3)) {
    set_error("vtu writer: arrays beyond 4 GiB (the format's 32-bit block sizes)");
    return SMG_ERR_ARG;
  }
  return SMG_OK;
}

}  // namespace
}  // namespace smg

using smg::set_error;

int smg_write_particles_vtu(const char* dir, int n_out, double time, const smg_vtu_points* p) {
  int rc = smg::check_points(p);
  if (rc != SMG_OK) return rc;
  if (n_out < 0) { set_error("smg_write_particles_vtu: n_out < 0"); return SMG_ERR_ARG; }
  const int64_t n = p->n;
  const std::string vtu = "particle_rank_0_step_" + std::to_string(n_out) + ".vtu";
  const std::string pvtu = "particle_step_" + std::to_string(n_out) + ".pvtu";
  std::ostringstream x;
  smg::Appended d;
  x << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n <UnstructuredGrid>\n"
    << "  <Piece NumberOfPoints=\"" << n << "\" NumberOfCells=\"" << n << "\">\n    <PointData Vectors=\"vector\">\n";
  std::vector<int32_t> ids(n);
  for (int64_t i = 0; i < n; i++) ids[i] = int32_t(i);
  d.add(x, "Int32", "id", 1, ids);
  d.add(x, "Int32", "local_id", 1, ids);   // one rank: the local index is the id
  d.add(x, "Int32", "type", 1, smg::or_zero(p->type, n, 1));
  d.add(x, "Float32", "h", 1, smg::to_float(p->h, n, 1));
  d.add(x, "Int32", "NN", 1, smg::or_zero(p->nn, n, 1));
  d.add(x, "Int32", "color", 1, smg::or_zero(p->color, n, 1));
  d.add(x, "Float32", "norm", 3, smg::or_zero(p->normal, n, 3));
  x << "    </PointData>\n    <Points>\n";
  d.add(x, "Float32", "Position", 3, smg::to_float(p->x, n, 3));
  x << "    </Points>\n    <Cells>\n";
  std::vector<int32_t> offs(n);
  std::vector<uint8_t> types(n, 1);   // VTK_VERTEX
  for (int64_t i = 0; i < n; i++) offs[i] = int32_t(i + 1);
  d.add(x, "Int32", "connectivity", 1, ids, false);
  d.add(x, "Int32", "offsets", 1, offs, false);
  d.add(x, "UInt8", "types", 1, types, false);
  x << "    </Cells>\n  </Piece>\n </UnstructuredGrid>\n <AppendedData encoding=\"raw\">\n _" << d.str()
    << "\n </AppendedData>\n</VTKFile>\n";
  rc = smg::write_file(smg::join(dir, vtu), x.str());
  if (rc != SMG_OK) return rc;
  rc = smg::write_pvtu(smg::join(dir, pvtu),
                       {{"Int32", "id", 1}, {"Int32", "local_id", 1}, {"Int32", "type", 1}, {"Float32", "h", 1},
                        {"Int32", "NN", 1}, {"Int32", "color", 1}, {"Float32", "norm", 3}},
                       {}, vtu);
  if (rc != SMG_OK) return rc;
  return smg::update_pvd(smg::join(dir, "particle.pvd"), n_out, time, pvtu);
}

int smg_write_mesh_vtu(const char* dir, int n_out, double time, int dim, const smg_vtu_points* p,
                       const int32_t* cells, int64_t n_cells) {
  int rc = smg::check_points(p);
  if (rc != SMG_OK) return rc;
  if (dim != 2 && dim != 3) { set_error("smg_write_mesh_vtu: dim must be 2 or 3"); return SMG_ERR_ARG; }
  if (n_out < 0 || n_cells < 0 || (n_cells > 0 && !cells)) { set_error("smg_write_mesh_vtu: bad cells"); return SMG_ERR_ARG; }
  const int per = dim == 2 ? 3 : 4;
  if (smg::too_big(n_cells, per)) { set_error("smg_write_mesh_vtu: arrays beyond 4 GiB"); return SMG_ERR_ARG; }
  const int64_t n = p->n;
  for (int64_t q = 0; q < n_cells * per; q++)
    if (cells[q] < 0 || cells[q] >= n) { set_error("smg_write_mesh_vtu: cell vertex out of range"); return SMG_ERR_ARG; }
  const std::string kind = dim == 2 ? "tris" : "tets";
  const std::string vtu = "mesh_" + kind + "_rank_0_step_" + std::to_string(n_out) + ".vtu";
  const std::string pvtu = "mesh_" + kind + "_step_" + std::to_string(n_out) + ".pvtu";
  std::ostringstream x;
  smg::Appended d;
  x << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n <UnstructuredGrid>\n"
    << "  <Piece NumberOfPoints=\"" << n << "\" NumberOfCells=\"" << n_cells << "\">\n"
    << "    <PointData Vectors=\"vector\">\n";
  d.add(x, "Int32", "type", 1, smg::or_zero(p->type, n, 1));
  d.add(x, "Float32", "h", 1, smg::to_float(p->h, n, 1));
  d.add(x, "Int32", "color", 1, smg::or_zero(p->color, n, 1));
  x << "    </PointData>\n    <CellData Vectors=\"vector\">\n";
  std::vector<smg::PointArrays> cell_arrays;
  if (dim == 3) {
    std::vector<float> rr(n_cells), dm(n_cells);
    smg::tet_cell_quality(p->x, cells, n_cells, rr.data(), dm.data());
    d.add(x, "Float32", "radius_ratio", 1, rr);
    d.add(x, "Float32", "mindihedangle", 1, dm);
    cell_arrays = {{"Float32", "radius_ratio", 1}, {"Float32", "mindihedangle", 1}};
  }
  x << "    </CellData>\n    <Points>\n";
  d.add(x, "Float32", "Position", 3, smg::to_float(p->x, n, 3));
  x << "    </Points>\n    <Cells>\n";
  std::vector<int32_t> conn(cells, cells + n_cells * per), offs(n_cells);
  for (int64_t q = 0; q < n_cells; q++) offs[q] = int32_t(per * (q + 1));
  std::vector<uint8_t> types(n_cells, dim == 2 ? 5 : 10);   // VTK_TRIANGLE, VTK_TETRA
  d.add(x, "Int32", "connectivity", 1, conn, false);
  d.add(x, "Int32", "offsets", 1, offs, false);
  d.add(x, "UInt8", "types", 1, types, false);
  x << "    </Cells>\n  </Piece>\n </UnstructuredGrid>\n <AppendedData encoding=\"raw\">\n _" << d.str()
    << "\n </AppendedData>\n</VTKFile>\n";
  rc = smg::write_file(smg::join(dir, vtu), x.str());
  if (rc != SMG_OK) return rc;
  rc = smg::write_pvtu(smg::join(dir, pvtu), {{"Int32", "type", 1}, {"Float32", "h", 1}, {"Int32", "color", 1}},
                       cell_arrays, vtu);
  if (rc != SMG_OK) return rc;
  return smg::update_pvd(smg::join(dir, "mesh_" + kind + ".pvd"), n_out, time, pvtu);
}

int smg_write_post_file(const char* dir, int64_t step, int64_t n, const double* x, int first) {
  if (n < 0 || n > 0x7fffffff || (n > 0 && !x)) { set_error("smg_write_post_file: bad arguments"); return SMG_ERR_ARG; }
  {
    const std::string list = smg::join(dir, "list.csv");
    std::ofstream f(list, first ? std::ios::trunc : std::ios::app);
    if (!f) { set_error("cannot write " + list); return SMG_ERR_ARG; }
    if (first) f << 1 << "\n";   // one rank
    f << step << " " << n << "\n";
    if (!f) { set_error("write failed: " + list); return SMG_ERR_ARG; }
  }
  const std::string pst = smg::join(dir, "step_" + std::to_string(step) + "_rank_0.pst");
  std::ofstream f(pst, std::ios::binary | std::ios::trunc);
  if (!f) { set_error("cannot write " + pst); return SMG_ERR_ARG; }
  const int32_t cnt = int32_t(n);
  f.write(reinterpret_cast<const char*>(&cnt), sizeof(cnt));
  if (n) f.write(reinterpret_cast<const char*>(x), sizeof(double) * 3 * n);
  if (!f) { set_error("write failed: " + pst); return SMG_ERR_ARG; }
  return SMG_OK;
}

int smg_read_post_list(const char* dir, int32_t* n_ranks, int64_t* steps, int64_t* counts, int64_t capacity,
                       int64_t* n_steps) {
  if (!n_ranks || !n_steps) { set_error("smg_read_post_list: NULL output"); return SMG_ERR_ARG; }
  const std::string list = smg::join(dir, "list.csv");
  std::ifstream f(list);
  if (!f) { set_error("cannot open " + list); return SMG_ERR_ARG; }
  long long r = 0;
  if (!(f >> r) || r <= 0 || r > 1000000) { set_error("list.csv: bad number of ranks"); return SMG_ERR_ARG; }
  *n_ranks = int32_t(r);
  int64_t k = 0;
  long long st = 0, np = 0;
  while (f >> st >> np) {
    if (np < 0) { set_error("list.csv: negative particle count"); return SMG_ERR_ARG; }
    if (steps && counts && k < capacity) {
      steps[k] = st;
      counts[k] = np;
    }
    k++;
  }
  if (!f.eof()) { set_error("list.csv: malformed line"); return SMG_ERR_ARG; }
  *n_steps = k;
  if ((steps || counts) && k > capacity) { set_error("smg_read_post_list: capacity"); return SMG_ERR_CAPACITY; }
  return SMG_OK;
}

int smg_read_post_file(const char* dir, int64_t step, int32_t n_ranks, double* x, int32_t* color, int64_t capacity,
                       int64_t* n) {
  if (!n || n_ranks <= 0) { set_error("smg_read_post_file: bad arguments"); return SMG_ERR_ARG; }
  int64_t total = 0;
  for (int32_t r = 0; r < n_ranks; r++) {
    const std::string pst = smg::join(dir, "step_" + std::to_string(step) + "_rank_" + std::to_string(r) + ".pst");
    std::ifstream f(pst, std::ios::binary);
    if (!f) { set_error("cannot open " + pst); return SMG_ERR_ARG; }
    int32_t cnt = 0;
    f.read(reinterpret_cast<char*>(&cnt), sizeof(cnt));
    if (!f || cnt < 0) { set_error("bad post file " + pst); return SMG_ERR_ARG; }
    if (x && total + cnt <= capacity) {
      f.read(reinterpret_cast<char*>(x + 3 * total), sizeof(double) * 3 * size_t(cnt));
      if (!f) { set_error("post file shorter than its count: " + pst); return SMG_ERR_ARG; }
      if (color)
        for (int64_t i = 0; i < cnt; i++) color[total + i] = r;   // Load_data_for_post: color = rank
    }
    total += cnt;
  }
  *n = total;
  if (x && total > capacity) { set_error("smg_read_post_file: capacity"); return SMG_ERR_CAPACITY; }
  return SMG_OK;
}
