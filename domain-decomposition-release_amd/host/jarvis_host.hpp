// Host-side C++ interface of the elastic FDTD / RTM path, with the names,
// argument meaning and error behaviour of the reference's host API, so that
// a driver written against the reference keeps compiling and reading the
// same project directories.  Everything here is plain C++ (g++); device
// work goes through the C ABI of include/ddx.h.
//
//   GeoConst / geo_const      BASE/fdtd3d_base_1_interface.hpp:10-17
//   SimulateType              BASE/fdtd3d_base_1_interface.hpp:19-26
//   ArgList                   BASE/fdtd3d_base_1_interface.hpp:27-113
//   Frame                     FLD/jarvis_0_frame_bones.hpp:9-47, _meat.hpp
//   Field<float>::read_raw / save / load
//                             FLD/jarvis_1_host_field_meat.hpp:20-197
//   GmsFileHeader / GmsReader UTIL/.jarvis_0_gms/jarvis_device_gms_bones.hpp:6-52
//   Geometry / SubGeometry    UTIL/.jarvis_2_geometry/jarvis_device_geometry_meat.hpp:6-128
//
// Errors keep the reference convention: a message on stdout and
// std::abort() (unreadable argfile, model, shot or receiver file; reading
// past the last GMS parameter).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace jarvis {

using std::string;

class GeoConst {
 public:
  static const uint32_t phy_fdorder_half = 4;
  static const uint32_t pml_fdorder_half = 2;
  static const uint32_t rec_width = 1;
};
using geo_const = GeoConst;

enum class SimulateType { pure_forward, rtm_forward, rtm_backward, rtm_reverse, fwi };

// <proj_path>argfile.json.  Keys and path prefixing as the reference
// (proj_path is prepended verbatim to model_path, shot_path, rece_path and
// output_path, so it normally ends with '/').  Keys this engine adds, all
// optional: rec_width (VOWR ring width, default GeoConst::rec_width),
// image_fp64 (0/1), source_mode (0 explosive, 1 szz only), cfl_policy
// (0 refuse an unstable model, 1 warn), device (HIP ordinal), image_output
// (0: image_pp and image_pp_laplace, 1: also the ps/sp/ss images, their
// Laplacians and the illumination ip), smooth_xyz [rx,ry,rz] (migrate with
// the model box-smoothed by these half-widths; modelling keeps the model),
// storage ("fp32", "fp16" or "bf16": wavefield storage format), ring_storage
// ("fp16" or "bf16": VOWR boundary store format over FP32 wavefields), and for raw
// models vp_path / vs_path / rho_path with n_xyz [nx,ny,nz] and d_xyz
// [dx,dy,dz].
struct ArgList {
  string proj_path;
  string model_path;
  string shot_path;
  string recv_path;
  int mpi_rank = 0;
  int mpi_size = 1;
  int x_divide_num = 1;
  int y_divide_num = 1;
  int z_divide_num = 1;
  int slots_per_node = 1;
  int pml_num = 0;
  float fm = 0.f;
  float T = 0.f;
  float dt = 0.f;
  bool is_output = false;
  string output_path;
  int t_snap = -1;
  // engine keys
  int rec_width = (int)GeoConst::rec_width;
  int image_fp64 = 0;
  int source_mode = 0;
  int cfl_policy = 0;
  int device = 0;
  int image_output = 0;  // 0: image_pp and its Laplacian (as the reference); 1: every image
  int smooth_xyz[3] = {0, 0, 0};  // > 0: migrate with the model box-smoothed by these half-widths
  int storage = 0;                // wavefield storage: 0 fp32, 1 fp16, 2 bf16
  int ring_storage = 0;           // VOWR store over FP32 wavefields: 0 as storage, 1 fp16, 2 bf16
  string vp_path, vs_path, rho_path;
  int n_xyz[3] = {0, 0, 0};
  float d_xyz[3] = {0.f, 0.f, 0.f};

  void mpi_read_arg_file(string _proj_path, int _mpi_rank, int _mpi_size, bool verbose = true);
};

// 72-byte frame header of the .fld format (field order and types as the
// reference struct, so .fld files are interchangeable).
struct Frame {
  uint64_t n_elem = 0, n_rows = 0, n_elem_slice = 0;
  uint32_t n_cols = 0, n_slices = 0;
  float d_rows = 0, d_cols = 0, d_slices = 0;
  float l_rows = 0, l_cols = 0, l_slices = 0;
  float r_rows = 0, r_cols = 0, r_slices = 0;
  enum class Type : int32_t { null = 0, vec = 1, mat = 2, cube = 3, line = 10, mesh = 20, grid = 30 };
  Type type = Type::null;

  Frame() {}
  Frame(uint64_t n_rows, uint32_t n_cols = 1, uint32_t n_slices = 1);
  Frame(uint64_t n_rows, uint32_t n_cols, uint32_t n_slices, float d_rows, float d_cols, float d_slices,
        float l_rows, float l_cols, float l_slices);
  void set_ndl(uint64_t n_rows, uint32_t n_cols, uint32_t n_slices, float d_rows, float d_cols, float d_slices,
               float l_rows, float l_cols, float l_slices);
  // "_nx_ny_nz_lx_ly_lz" with the origins printed as std::to_string(float)
  string size_info() const;

 private:
  void set_frame_type();
};
static_assert(sizeof(Frame) == 72, "Frame must match the 72-byte .fld header");

enum class SaveFormat { binary_raw, binary_fld, ascii_xyz, ascii_grd, ascii_txt };

// Host field, x (rows) fastest.
template <typename T>
struct Field : public Frame {
  std::vector<T> data;
  void alloc(const Frame &f) {
    static_cast<Frame &>(*this) = f;
    data.assign(n_elem, T(0));
  }
  T *ptr() { return data.data(); }
  const T *ptr() const { return data.data(); }
  T &operator()(uint64_t i, uint32_t j = 0, uint32_t k = 0) { return data[i + j * n_rows + k * n_elem_slice]; }
  T operator()(uint64_t i, uint32_t j = 0, uint32_t k = 0) const { return data[i + j * n_rows + k * n_elem_slice]; }
  // raw float data, no header (FLD/jarvis_1_host_field_meat.hpp:65-79)
  void read_raw(string path, const Frame &frame);
  // binary_raw appends size_info() + ".raw", binary_fld ".fld", ascii_txt
  // ".txt" (one line per column, "%10.6e\t" per value, for 1D/2D frames)
  void save(string path, SaveFormat format) const;
  // .fld: 72-byte Frame then the data (FLD/jarvis_1_host_field_meat.hpp:42-63)
  void load(string path);
  void clear() {
    data.clear();
    data.shrink_to_fit();
  }
};

// GMS model file: 240-byte header, then para_num blocks of a 64-byte name
// and n_elem floats stored z fastest, then x, then y.  The header struct uses
// natural alignment: para_num at byte 74, l_* at 76, n_* at 88, d_* at 100.
struct GmsFileHeader {
  int16_t format;
  char name[64];
  int16_t year;
  char mouth;
  char day;
  char hour;
  char minutes;
  char seconds;
  char type;
  int16_t para_num;
  float l_rows, l_cols, l_slices;
  uint32_t n_rows, n_cols, n_slices;
  float d_rows, d_cols, d_slices;
  char reserved[127];
};
static_assert(sizeof(GmsFileHeader) == 240, "GMS header is 240 bytes");

class GmsReader {
 public:
  Frame grid;
  GmsFileHeader file_header{};
  std::vector<string> names;
  std::vector<std::vector<float>> data;
  explicit GmsReader(std::string path);
  // next parameter (vp, vs, rho in order) into an x-fastest field
  void read_gms_by_order_to_ffield(Field<float> &field_obj);
  void print_info() const;
  void clear();

 private:
  int para_order = 0;
  int para_num = 0;
  std::string gms_model_path;
  void read_gms_data();
};

// Writes a GMS file (no counterpart in the reference, which only reads);
// used to produce model files and by the format tests.
void write_gms(const string &path, const Frame &grid, const std::vector<string> &names,
               const std::vector<const float *> &fields_xfast);

struct Point3 {
  float x = 0.f, y = 0.f, z = 0.f;
  int ind = 0;
};

// Shot / receiver text files: a count, then x y z triples in metres.
class Geometry {
 public:
  string shot_path, recv_path;
  std::vector<Point3> shot, recv;
  void read_shot_line();
  void read_recv_line();
};

}  // namespace jarvis
