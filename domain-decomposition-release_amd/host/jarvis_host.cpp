#include "jarvis_host.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "json_lite.hpp"

namespace jarvis {

[[noreturn]] static void die(const string &msg) {
  std::cout << msg << std::endl;
  std::abort();
}

// ---------------------------------------------------------------------------
// ArgList (BASE/fdtd3d_base_1_interface.hpp:50-112)
// ---------------------------------------------------------------------------
void ArgList::mpi_read_arg_file(string _proj_path, int _mpi_rank, int _mpi_size, bool verbose) {
  std::ifstream fin(_proj_path + "argfile.json");
  if (!fin.is_open()) die("\033[41;37marg_file open error!\033[0m");
  std::stringstream ss;
  ss << fin.rdbuf();
  ddxhost::JsonValue value;
  std::string err;
  // The reference ignores a parse failure and reads whatever parsed (missing
  // keys read as 0 / "" / false); so does this reader, with a warning.
  if (!ddxhost::json_parse(ss.str(), value, err))
    std::cerr << "argfile.json: " << err << " (continuing with the keys read so far)" << std::endl;
  try {
    model_path = _proj_path + value["model_path"].asString();
    shot_path = _proj_path + value["shot_path"].asString();
    recv_path = _proj_path + value["rece_path"].asString();
    x_divide_num = value["x_divide_num"].asInt();
    y_divide_num = value["y_divide_num"].asInt();
    z_divide_num = value["z_divide_num"].asInt();
    slots_per_node = value["slots_per_node"].asInt();
    pml_num = value["pml_num"].asInt();
    fm = value["fm"].asFloat();
    T = value["T"].asFloat();
    dt = value["dt"].asFloat();
    is_output = value["is_output"].asBool();
    output_path = _proj_path + value["output_path"].asString();
    t_snap = value["t_snap"].asInt();
    // engine keys (absent in the reference's argfiles)
    if (!value["rec_width"].is_null()) rec_width = value["rec_width"].asInt();
    image_fp64 = value["image_fp64"].asInt();
    source_mode = value["source_mode"].asInt();
    cfl_policy = value["cfl_policy"].asInt();
    device = value["device"].asInt();
    image_output = value["image_output"].asInt();
    for (const char *key : {"storage", "ring_storage"}) {
      if (value[key].is_null()) continue;
      const string st = value[key].asString();
      int id;
      if (st == "fp32" || st == "0") id = 0;
      else if (st == "fp16" || st == "1") id = 1;
      else if (st == "bf16" || st == "2") id = 2;
      else if (string(key) == "ring_storage" && (st == "bfp16" || st == "3")) id = 3;
      else die(string("argfile.json: ") + key + " must be \"fp32\", \"fp16\" or \"bf16\"" +
               (string(key) == "ring_storage" ? " (or \"bfp16\")" : ""));
      (string(key) == "storage" ? storage : ring_storage) = id;
    }
    if (!value["vp_path"].is_null()) vp_path = _proj_path + value["vp_path"].asString();
    if (!value["vs_path"].is_null()) vs_path = _proj_path + value["vs_path"].asString();
    if (!value["rho_path"].is_null()) rho_path = _proj_path + value["rho_path"].asString();
    for (int a = 0; a < 3; a++) {
      n_xyz[a] = value["n_xyz"][a].asInt();
      d_xyz[a] = value["d_xyz"][a].asFloat();
      smooth_xyz[a] = value["smooth_xyz"][a].asInt();
    }
  } catch (const std::exception &e) {
    die(string("argfile.json: ") + e.what());
  }
  proj_path = _proj_path;
  mpi_rank = _mpi_rank;
  mpi_size = _mpi_size;
  if (verbose && mpi_rank == 0) {
    std::cout << "-------------------------" << std::endl;
    std::cout << "proj_path:\t" << proj_path << std::endl;
    std::cout << "model_path:\t" << model_path << std::endl;
    std::cout << "shot_path:\t" << shot_path << std::endl;
    std::cout << "recv_path:\t" << recv_path << std::endl;
    std::cout << "x_divide_num:\t" << x_divide_num << std::endl;
    std::cout << "y_divide_num:\t" << y_divide_num << std::endl;
    std::cout << "z_divide_num:\t" << z_divide_num << std::endl;
    std::cout << "mpi_size:\t" << mpi_size << std::endl;
    std::cout << "slots_per_node:\t" << slots_per_node << std::endl;
    std::cout << "pml_num:\t" << pml_num << std::endl;
    std::cout << "fm:     \t" << fm << std::endl;
    std::cout << "dt:     \t" << dt << std::endl;
    std::cout << "T:      \t" << T << std::endl;
    std::cout << "is_output:\t" << is_output << std::endl;
    std::cout << "output_path:\t" << output_path << std::endl;
    std::cout << "rec_width:\t" << rec_width << std::endl;
    std::cout << "-------------------------" << std::endl;
  }
}

// ---------------------------------------------------------------------------
// Frame (FLD/jarvis_0_frame_meat.hpp)
// ---------------------------------------------------------------------------
Frame::Frame(uint64_t nr, uint32_t nc, uint32_t ns) { set_ndl(nr, nc, ns, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f); }

Frame::Frame(uint64_t nr, uint32_t nc, uint32_t ns, float dr, float dc, float ds, float lr, float lc, float ls) {
  set_ndl(nr, nc, ns, dr, dc, ds, lr, lc, ls);
}

void Frame::set_ndl(uint64_t nr, uint32_t nc, uint32_t ns, float dr, float dc, float ds, float lr, float lc,
                    float ls) {
  n_rows = nr;
  n_cols = nc;
  n_slices = ns;
  n_elem_slice = n_rows * n_cols;
  n_elem = n_elem_slice * n_slices;
  d_rows = dr;
  d_cols = dc;
  d_slices = ds;
  l_rows = lr;
  l_cols = lc;
  l_slices = ls;
  // right edges in float, (n-1)*d added to l; an empty axis has l = r = 0
  if (n_rows > 0) r_rows = l_rows + (float)(n_rows - 1) * d_rows; else l_rows = r_rows = 0;
  if (n_cols > 0) r_cols = l_cols + (float)(n_cols - 1) * d_cols; else l_cols = r_cols = 0;
  if (n_slices > 0) r_slices = l_slices + (float)(n_slices - 1) * d_slices; else l_slices = r_slices = 0;
  set_frame_type();
}

// Type from the sizes and whether any spacing is set (FLD/jarvis_0_frame_meat.hpp:219-239).
void Frame::set_frame_type() {
  const bool one_d = (n_rows == 1 && n_cols == 1) || (n_rows == 1 && n_slices == 1) || (n_cols == 1 && n_slices == 1);
  const bool two_d = (n_rows > 1 && n_cols > 1 && n_slices == 1) || (n_rows > 1 && n_slices > 1 && n_cols == 1) ||
                     (n_cols > 1 && n_slices > 1 && n_rows == 1);
  const bool three_d = n_rows > 1 && n_cols > 1 && n_slices > 1;
  const bool spaced = d_rows > 0 || d_cols > 0 || d_slices > 0;
  if (one_d) type = spaced ? Type::line : Type::vec;
  else if (two_d) type = spaced ? Type::mesh : Type::mat;
  else if (three_d) type = spaced ? Type::grid : Type::cube;
}

string Frame::size_info() const {
  return "_" + std::to_string(n_rows) + "_" + std::to_string(n_cols) + "_" + std::to_string(n_slices) + "_" +
         std::to_string(l_rows) + "_" + std::to_string(l_cols) + "_" + std::to_string(l_slices);
}

// ---------------------------------------------------------------------------
// Field<float> I/O (FLD/jarvis_1_host_field_meat.hpp)
// ---------------------------------------------------------------------------
template <typename T>
void Field<T>::read_raw(string path, const Frame &frame) {
  FILE *fp = fopen(path.c_str(), "rb");
  if (!fp) die("load():\033[41;37m[error]:\033[0m File open error!");
  alloc(frame);
  size_t got = fread(data.data(), sizeof(T), n_elem, fp);
  fclose(fp);
  if (got != n_elem)
    std::cerr << "read_raw: " << path << " holds " << got << " of " << n_elem << " values; the rest read as 0"
              << std::endl;
}

template <typename T>
void Field<T>::save(string path, SaveFormat format) const {
  FILE *fp = nullptr;
  switch (format) {
    case SaveFormat::binary_raw:
      path += size_info() + ".raw";
      fp = fopen(path.c_str(), "wb");
      if (!fp) die("save():\033[41;37m[error]:\033[0m can't create " + path);
      fwrite(data.data(), sizeof(T), n_elem, fp);
      break;
    case SaveFormat::binary_fld:
      path += ".fld";
      fp = fopen(path.c_str(), "wb");
      if (!fp) die("save():\033[41;37m[error]:\033[0m can't create " + path);
      fwrite(static_cast<const Frame *>(this), sizeof(Frame), 1, fp);
      fwrite(data.data(), sizeof(T), n_elem, fp);
      break;
    case SaveFormat::ascii_txt:
      path += ".txt";
      fp = fopen(path.c_str(), "wt");
      if (!fp) die("save_as_ascii_txt():\033[41;37m[error]:\033[0mcan't create txt output file " + path);
      if (type == Type::null) die("save_as_ascii_txt():\033[41;37m[error]:\033[0mframe type is null!");
      if (type == Type::cube || type == Type::grid) {
        for (uint32_t k = 0; k < n_slices; ++k)
          for (uint32_t j = 0; j < n_cols; ++j)
            for (uint64_t i = 0; i < n_rows; ++i)
              fprintf(fp, "%d,%d,%d,%10.6e\n", (int)i, (int)j, (int)k, (double)(*this)(i, j, k));
      } else {
        for (uint32_t j = 0; j < n_cols; ++j) {
          for (uint64_t i = 0; i < n_rows; ++i) fprintf(fp, "%10.6e\t", (double)(*this)(i, j));
          fprintf(fp, "\n");
        }
      }
      break;
    default:
      die("save(): only binary_raw, binary_fld and ascii_txt are provided");
  }
  fclose(fp);
}

template <typename T>
void Field<T>::load(string path) {
  if (path.size() < 4 || path.substr(path.size() - 4) != ".fld")
    die("load:\033[41;37m[error]:\033[0m " + path + " Can't read field-raw file %s ,exit");
  FILE *fp = fopen(path.c_str(), "rb");
  if (!fp) die("load():\033[41;37m[error]:\033[0m File open error!");
  Frame f;
  if (fread(&f, sizeof(Frame), 1, fp) != 1) {
    fclose(fp);
    die("load(): " + path + " is shorter than its 72-byte header");
  }
  alloc(f);
  size_t got = fread(data.data(), sizeof(T), n_elem, fp);
  fclose(fp);
  if (got != n_elem) die("load(): " + path + " holds fewer values than its header says");
}

template struct Field<float>;

// ---------------------------------------------------------------------------
// GMS (UTIL/.jarvis_0_gms/jarvis_device_gms_meat.hpp:11-59)
// ---------------------------------------------------------------------------
GmsReader::GmsReader(std::string path) : gms_model_path(std::move(path)) { read_gms_data(); }

void GmsReader::read_gms_data() {
  FILE *fp = fopen(gms_model_path.c_str(), "rb");
  if (!fp) {
    printf("read_gms_data():\033[41;37m[error]:\033[0mFile open error!");
    fflush(stdout);
    std::abort();
  }
  if (fread(&file_header, 240, 1, fp) != 1) {
    fclose(fp);
    die("read_gms_data(): " + gms_model_path + " is shorter than the 240-byte header");
  }
  grid.set_ndl(file_header.n_rows, file_header.n_cols, file_header.n_slices, file_header.d_rows, file_header.d_cols,
               file_header.d_slices, file_header.l_rows, file_header.l_cols, file_header.l_slices);
  para_num = file_header.para_num;
  names.assign(para_num, "");
  data.assign(para_num, {});
  for (int p = 0; p < para_num; p++) {
    char nm[65] = {0};
    data[p].assign(grid.n_elem, 0.f);
    size_t a = fread(nm, 64, 1, fp);
    size_t b = fread(data[p].data(), sizeof(float), grid.n_elem, fp);
    names[p] = nm;
    if (a != 1 || b != grid.n_elem) {
      fclose(fp);
      die("read_gms_data(): " + gms_model_path + " ends inside parameter " + std::to_string(p));
    }
  }
  fclose(fp);
}

void GmsReader::read_gms_by_order_to_ffield(Field<float> &f) {
  if (para_order >= para_num) {
    printf("ERROR:There are no more parameters to read!");
    fflush(stdout);
    std::abort();
  }
  f.alloc(grid);
  const uint64_t nr = grid.n_rows, nc = grid.n_cols, ns = grid.n_slices;
  const std::vector<float> &src = data[para_order];
  for (uint64_t k = 0; k < ns; k++)
    for (uint64_t j = 0; j < nc; j++)
      for (uint64_t i = 0; i < nr; i++) f(i, j, k) = src[k + i * ns + j * nr * ns];
  para_order++;
}

void GmsReader::print_info() const {
  std::cout << "Grid Information:" << std::endl;
  std::cout << "para_order:" << para_order << ";" << "para_num:" << para_num << ";" << std::endl;
  std::cout << "GMS_MODEL: n " << grid.n_rows << " " << grid.n_cols << " " << grid.n_slices << " d " << grid.d_rows
            << " " << grid.d_cols << " " << grid.d_slices << " l " << grid.l_rows << " " << grid.l_cols << " "
            << grid.l_slices << std::endl;
}

void GmsReader::clear() {
  data.clear();
  data.shrink_to_fit();
}

void write_gms(const string &path, const Frame &grid, const std::vector<string> &names,
               const std::vector<const float *> &fields) {
  GmsFileHeader h;
  memset(&h, 0, sizeof h);
  h.para_num = (int16_t)fields.size();
  h.l_rows = grid.l_rows;
  h.l_cols = grid.l_cols;
  h.l_slices = grid.l_slices;
  h.n_rows = (uint32_t)grid.n_rows;
  h.n_cols = grid.n_cols;
  h.n_slices = grid.n_slices;
  h.d_rows = grid.d_rows;
  h.d_cols = grid.d_cols;
  h.d_slices = grid.d_slices;
  FILE *fp = fopen(path.c_str(), "wb");
  if (!fp) die("write_gms: can't create " + path);
  fwrite(&h, sizeof h, 1, fp);
  const uint64_t nr = grid.n_rows, nc = grid.n_cols, ns = grid.n_slices;
  std::vector<float> buf(grid.n_elem);
  for (size_t p = 0; p < fields.size(); p++) {
    char nm[64] = {0};
    if (p < names.size()) strncpy(nm, names[p].c_str(), 63);
    fwrite(nm, 64, 1, fp);
    for (uint64_t k = 0; k < ns; k++)
      for (uint64_t j = 0; j < nc; j++)
        for (uint64_t i = 0; i < nr; i++) buf[k + i * ns + j * nr * ns] = fields[p][i + j * nr + k * nr * nc];
    fwrite(buf.data(), sizeof(float), buf.size(), fp);
  }
  fclose(fp);
}

// ---------------------------------------------------------------------------
// Shot and receiver lines (UTIL/.jarvis_2_geometry/jarvis_device_geometry_meat.hpp:6-55)
// ---------------------------------------------------------------------------
static void read_points(const string &path, std::vector<Point3> &out, const char *who) {
  std::ifstream in(path);
  if (!in.is_open()) {
    printf("%s(path):file open error!\n", who);
    fflush(stdout);
    std::abort();
  }
  long n = 0;
  in >> n;
  if (n < 0) n = 0;
  out.assign((size_t)n, Point3());
  for (long i = 0; i < n; i++) {
    in >> out[i].x >> out[i].y >> out[i].z;
    out[i].ind = (int)i;
  }
}

void Geometry::read_shot_line() { read_points(shot_path, shot, "read_shot_line"); }
void Geometry::read_recv_line() { read_points(recv_path, recv, "read_recv_line"); }

}  // namespace jarvis
