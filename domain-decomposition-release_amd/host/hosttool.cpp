// Command-line access to the host readers and writers, for the format tests
// (no GPU, no engine):
//   ddx_hosttool argfile <proj_path>                  ArgList as JSON
//   ddx_hosttool points <file>                        shot/receiver line as JSON
//   ddx_hosttool gms2raw <file.gms> <out_prefix>      each parameter x-fastest, Field::save(binary_raw)
//   ddx_hosttool writegms <out.gms> nx ny nz dx dy dz lx ly lz <raw>...   GMS from x-fastest raw files
//   ddx_hosttool txt <raw> <n_rows> <n_cols> <out_prefix>                Field::save(ascii_txt)
//   ddx_hosttool fld <raw> nx ny nz dx dy dz lx ly lz <out_prefix>      Field::save(binary_fld), load, compare
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "jarvis_host.hpp"

using namespace jarvis;

static void frame_json(const Frame &f) {
  printf("{\"n\": [%llu, %u, %u], \"d\": [%.9g, %.9g, %.9g], \"l\": [%.9g, %.9g, %.9g], "
         "\"r\": [%.9g, %.9g, %.9g], \"type\": %d, \"size_info\": \"%s\"}",
         (unsigned long long)f.n_rows, f.n_cols, f.n_slices, f.d_rows, f.d_cols, f.d_slices, f.l_rows, f.l_cols,
         f.l_slices, f.r_rows, f.r_cols, f.r_slices, (int)f.type, f.size_info().c_str());
}

static Frame frame_args(char **a) {
  return Frame(strtoull(a[0], nullptr, 10), (uint32_t)atoi(a[1]), (uint32_t)atoi(a[2]), (float)atof(a[3]),
               (float)atof(a[4]), (float)atof(a[5]), (float)atof(a[6]), (float)atof(a[7]), (float)atof(a[8]));
}

static std::string esc(const std::string &s) {
  std::string o;
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    o += c;
  }
  return o;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: ddx_hosttool <argfile|points|gms2raw|writegms|txt|fld> ...\n");
    return 2;
  }
  const std::string cmd = argv[1];
  if (cmd == "argfile") {
    ArgList a;
    a.mpi_read_arg_file(argv[2], 0, 1, false);
    printf("{\"proj_path\": \"%s\", \"model_path\": \"%s\", \"shot_path\": \"%s\", \"recv_path\": \"%s\", "
           "\"output_path\": \"%s\", \"x_divide_num\": %d, \"y_divide_num\": %d, \"z_divide_num\": %d, "
           "\"slots_per_node\": %d, \"pml_num\": %d, \"fm\": %.9g, \"T\": %.9g, \"dt\": %.9g, \"is_output\": %d, "
           "\"t_snap\": %d, \"rec_width\": %d, \"image_fp64\": %d, \"source_mode\": %d, \"cfl_policy\": %d, "
           "\"device\": %d, \"image_output\": %d, \"storage\": %d, \"ring_storage\": %d, \"smooth_xyz\": [%d, %d, %d], \"n_xyz\": [%d, %d, %d], "
           "\"d_xyz\": [%.9g, %.9g, %.9g]}\n",
           esc(a.proj_path).c_str(), esc(a.model_path).c_str(), esc(a.shot_path).c_str(), esc(a.recv_path).c_str(),
           esc(a.output_path).c_str(), a.x_divide_num, a.y_divide_num, a.z_divide_num, a.slots_per_node, a.pml_num,
           a.fm, a.T, a.dt, (int)a.is_output, a.t_snap, a.rec_width, a.image_fp64, a.source_mode, a.cfl_policy,
           a.device, a.image_output, a.storage, a.ring_storage, a.smooth_xyz[0], a.smooth_xyz[1], a.smooth_xyz[2], a.n_xyz[0], a.n_xyz[1], a.n_xyz[2], a.d_xyz[0], a.d_xyz[1], a.d_xyz[2]);
    return 0;
  }
  if (cmd == "points") {
    Geometry g;
    g.shot_path = argv[2];
    g.read_shot_line();
    printf("[");
    for (size_t i = 0; i < g.shot.size(); i++)
      printf("%s[%.9g, %.9g, %.9g, %d]", i ? ", " : "", g.shot[i].x, g.shot[i].y, g.shot[i].z, g.shot[i].ind);
    printf("]\n");
    return 0;
  }
  if (cmd == "gms2raw" && argc == 4) {
    GmsReader gms(argv[2]);
    printf("{\"para_num\": %d, \"names\": [", (int)gms.names.size());
    for (size_t p = 0; p < gms.names.size(); p++) printf("%s\"%s\"", p ? ", " : "", esc(gms.names[p]).c_str());
    printf("], \"files\": [");
    for (size_t p = 0; p < gms.names.size(); p++) {
      Field<float> f;
      gms.read_gms_by_order_to_ffield(f);
      const std::string base = std::string(argv[3]) + std::to_string(p);
      f.save(base, SaveFormat::binary_raw);
      printf("%s\"%s\"", p ? ", " : "", esc(base + f.size_info() + ".raw").c_str());
    }
    printf("], \"frame\": ");
    frame_json(gms.grid);
    printf("}\n");
    return 0;
  }
  if (cmd == "writegms" && argc >= 13) {
    const Frame f = frame_args(argv + 3);
    std::vector<Field<float>> fields(argc - 12);
    std::vector<const float *> ptrs;
    std::vector<std::string> names;
    for (int p = 12; p < argc; p++) {
      fields[p - 12].read_raw(argv[p], f);
      ptrs.push_back(fields[p - 12].ptr());
      names.push_back("para" + std::to_string(p - 12));
    }
    write_gms(argv[2], f, names, ptrs);
    return 0;
  }
  if (cmd == "txt" && argc == 6) {
    Field<float> f;
    f.read_raw(argv[2], Frame(strtoull(argv[3], nullptr, 10), (uint32_t)atoi(argv[4])));
    f.save(argv[5], SaveFormat::ascii_txt);
    return 0;
  }
  if (cmd == "fld" && argc == 13) {
    const Frame fr = frame_args(argv + 3);
    Field<float> f, g;
    f.read_raw(argv[2], fr);
    f.save(argv[12], SaveFormat::binary_fld);
    g.load(std::string(argv[12]) + ".fld");
    const bool same = memcmp(static_cast<const Frame *>(&f), static_cast<const Frame *>(&g), sizeof(Frame)) == 0 &&
                      f.data == g.data;
    printf("{\"same\": %s, \"frame\": ", same ? "true" : "false");
    frame_json(g);
    printf("}\n");
    return same ? 0 : 1;
  }
  fprintf(stderr, "ddx_hosttool: bad command or argument count\n");
  return 2;
}
