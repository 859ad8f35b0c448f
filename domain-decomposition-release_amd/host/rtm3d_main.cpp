// RTM driver over the C ABI: the program the reference builds from
// RTMTEST/rtm3d_blend_elastic_model_mpi_main.cpp + _device.cu, with every
// subdomain of the decomposition on one GPU instead of one MPI rank each.
//
//   ddx_rtm3d <proj_path> [--forward-only]
//
// Reads <proj_path>argfile.json, the model (GMS, or three raw files), the
// shot and receiver lines; runs the modelling loop and writes
// <output_path>v{x,y,z}_seis_<rank>_<shot>.txt for every subdomain that holds
// receivers; then (unless --forward-only) the RTM loop over all shots (the
// images stack over shots, in double with image_fp64 = 1), the Laplacian and
// source compensation, and writes
// <output_path>image_pp_<rank>_<nx>_<ny>_<nz>_<lx>_<ly>_<lz>.raw and
// image_pp_laplace_<rank>_....raw per subdomain
// (RTM/rtm3d_blend_elastic_model_mpi_meat.hpp:18-135); image_output = 1
// adds the ps, sp and ss images, their Laplacians and image_ip.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "ddx.h"
#include "jarvis_host.hpp"

using namespace jarvis;

static void check(int rc, ddx_ctx *c, const char *what) {
  if (rc != DDX_OK) {
    std::cout << "\033[41;37m[error]\033[0m " << what << ": " << ddx_last_error(c) << std::endl;
    std::abort();
  }
}

static bool ends_with(const string &s, const string &e) {
  return s.size() >= e.size() && s.compare(s.size() - e.size(), e.size(), e) == 0;
}

// Origin of a subdomain's gridphy in metres, computed as the reference does
// (BASE/fdtd3d_base_3_geomodel_meat.hpp:579-640): 0 for the first subdomain
// of an axis, else i * int(ext / ndiv) * d - pml * d in float.
static float phy_origin(int i, int n, int ext, int pml, float d) {
  return i == 0 ? 0.f : (float)(i * (ext / n)) * d - (float)pml * d;
}

int main(int argc, char *argv[]) {
  if (argc < 2 || !argv[1]) {
    std::cout << "Please input argfile!" << std::endl;
    std::abort();
  }
  const bool forward_only = argc > 2 && string(argv[2]) == "--forward-only";
  const string proj_path = argv[1];
  ArgList arg_list;
  arg_list.mpi_read_arg_file(proj_path, 0, 1);

  // ---- model ----
  Field<float> vp, vs, rho;
  if (ends_with(arg_list.model_path, ".gms")) {
    GmsReader gms(arg_list.model_path);
    gms.read_gms_by_order_to_ffield(vp);
    gms.read_gms_by_order_to_ffield(vs);
    gms.read_gms_by_order_to_ffield(rho);
    gms.clear();
  } else {
    const Frame f(arg_list.n_xyz[0], arg_list.n_xyz[1], arg_list.n_xyz[2], arg_list.d_xyz[0], arg_list.d_xyz[1],
                  arg_list.d_xyz[2], 0.f, 0.f, 0.f);
    if (f.n_elem == 0) {
      std::cout << "model: model_path is not a .gms file and n_xyz is not set for raw models" << std::endl;
      std::abort();
    }
    vp.read_raw(arg_list.vp_path, f);
    vs.read_raw(arg_list.vs_path, f);
    rho.read_raw(arg_list.rho_path, f);
  }

  // ---- survey ----
  Geometry geo;
  geo.shot_path = arg_list.shot_path;
  geo.recv_path = arg_list.recv_path;
  geo.read_shot_line();
  geo.read_recv_line();
  std::vector<float> sxyz, rxyz;
  for (const Point3 &p : geo.shot) sxyz.insert(sxyz.end(), {p.x, p.y, p.z});
  for (const Point3 &p : geo.recv) rxyz.insert(rxyz.end(), {p.x, p.y, p.z});

  // ---- engine ----
  ddx_desc d{};
  d.nx = (int)vp.n_rows;
  d.ny = (int)vp.n_cols;
  d.nz = (int)vp.n_slices;
  d.dx = vp.d_rows;
  d.dy = vp.d_cols;
  d.dz = vp.d_slices;
  d.pml_num = arg_list.pml_num;
  d.x_divide_num = arg_list.x_divide_num;
  d.y_divide_num = arg_list.y_divide_num;
  d.z_divide_num = arg_list.z_divide_num;
  d.dt = arg_list.dt;
  d.fm = arg_list.fm;
  d.ntime = ddx_ntime(arg_list.T, arg_list.dt);
  d.rec_width = arg_list.rec_width;
  d.image_fp64 = arg_list.image_fp64;
  d.source_mode = arg_list.source_mode;
  d.device = arg_list.device;
  d.cfl_policy = arg_list.cfl_policy;
  d.storage = arg_list.storage;
  d.ring_storage = arg_list.ring_storage;
  ddx_ctx *ctx = nullptr;
  check(ddx_create(&d, &ctx), nullptr, "ddx_create");
  check(ddx_set_model(ctx, vp.ptr(), vs.ptr(), rho.ptr()), ctx, "ddx_set_model");
  const int *sm = arg_list.smooth_xyz;
  if (sm[0] > 0 || sm[1] > 0 || sm[2] > 0) {
    // migration background: the model box-smoothed (modelling keeps the model)
    for (Field<float> *f : {&vp, &vs, &rho})
      check(ddx_smooth_model(f->ptr(), f->ptr(), d.nx, d.ny, d.nz, sm[0], sm[1], sm[2]), ctx, "ddx_smooth_model");
    check(ddx_set_migration_model(ctx, vp.ptr(), vs.ptr(), rho.ptr()), ctx, "ddx_set_migration_model");
  }
  vp.clear();
  vs.clear();
  rho.clear();
  std::vector<float> wavelet(d.ntime);
  check(ddx_ricker(d.fm, d.dt, d.ntime, wavelet.data()), ctx, "ddx_ricker");
  check(ddx_set_wavelet(ctx, wavelet.data()), ctx, "ddx_set_wavelet");
  const int nshot = (int)geo.shot.size();
  check(ddx_set_geometry(ctx, sxyz.data(), nshot, rxyz.data(), (int)geo.recv.size()), ctx, "ddx_set_geometry");
  int nkept = 0;
  check(ddx_receivers(ctx, &nkept, nullptr, nullptr), ctx, "ddx_receivers");
  std::vector<int> rank_of_row(nkept), input_of_row(nkept);
  check(ddx_receivers(ctx, &nkept, rank_of_row.data(), input_of_row.data()), ctx, "ddx_receivers");
  const int nsub = d.x_divide_num * d.y_divide_num * d.z_divide_num;
  const int nt = d.ntime;

  // ---- modelling loop (RTM/...meat.hpp:20-39) ----
  std::vector<float> seis((size_t)3 * nkept * nt);
  auto t0 = std::chrono::steady_clock::now();
  for (int ishot = 0; ishot < nshot; ishot++) {
    printf("\033[44;37mshot: %d\033[0m\n", ishot + 1);
    check(ddx_model_shot(ctx, ishot), ctx, "ddx_model_shot");
    check(ddx_get_seismograms(ctx, ishot, seis.data()), ctx, "ddx_get_seismograms");
    for (int r = 0; r < nsub; r++) {
      std::vector<int> rows;
      for (int q = 0; q < nkept; q++)
        if (rank_of_row[q] == r) rows.push_back(q);
      if (rows.empty()) continue;
      static const char *comp[3] = {"vx", "vy", "vz"};
      for (int c = 0; c < 3; c++) {
        Field<float> f;
        f.alloc(Frame((uint64_t)nt, (uint32_t)rows.size()));
        for (size_t q = 0; q < rows.size(); q++)
          for (int it = 0; it < nt; it++) f(it, (uint32_t)q) = seis[((size_t)c * nkept + rows[q]) * nt + it];
        f.save(arg_list.output_path + comp[c] + "_seis_" + std::to_string(r) + "_" + std::to_string(ishot),
               SaveFormat::ascii_txt);
      }
    }
  }
  auto t1 = std::chrono::steady_clock::now();
  printf("modelling: %d shot(s) in %.3f s\n", nshot, std::chrono::duration<double>(t1 - t0).count());
  if (forward_only) {
    ddx_destroy(ctx);
    return 0;
  }

  // ---- RTM loop (RTM/...meat.hpp:50-122) ----
  for (int ishot = 0; ishot < nshot; ishot++) {
    printf("\033[44;37mshot: %d\033[0m\n", ishot + 1);
    auto a = std::chrono::steady_clock::now();
    check(ddx_rtm_shot(ctx, ishot), ctx, "ddx_rtm_shot");
    check(ddx_synchronize(ctx), ctx, "ddx_synchronize");
    auto b = std::chrono::steady_clock::now();
    printf("rtm of one shot time: %.3f s\n", std::chrono::duration<double>(b - a).count());
  }
  check(ddx_finalize_images(ctx), ctx, "ddx_finalize_images");

  // ---- image output per subdomain (RTM/...meat.hpp:126-134) ----
  const int nd[3] = {d.x_divide_num, d.y_divide_num, d.z_divide_num};
  const int gn[3] = {d.nx, d.ny, d.nz};
  const float dd[3] = {d.dx, d.dy, d.dz};
  for (int r = 0; r < nsub; r++) {
    int g[21];
    check(ddx_subdomain(&d, r, g), ctx, "ddx_subdomain");
    const int ci[3] = {r % (nd[0] * nd[1]) % nd[0], r % (nd[0] * nd[1]) / nd[0], r / (nd[0] * nd[1])};
    float l[3];
    for (int a = 0; a < 3; a++) l[a] = phy_origin(ci[a], nd[a], gn[a] + 2 * d.pml_num, d.pml_num, dd[a]);
    const Frame f(g[0], g[1], g[2], dd[0], dd[1], dd[2], l[0], l[1], l[2]);
    Field<float> img;
    img.alloc(f);
    // image_pp and image_pp_laplace as the reference; with image_output = 1
    // also the P-S, S-P and S-S images, their Laplacians and the source
    // illumination Ip
    static const struct {
      int id;
      const char *name;
    } outs[] = {{DDX_IPP, "image_pp_"},         {DDX_IPP_LAPLACE, "image_pp_laplace_"},
                {DDX_IPS, "image_ps_"},         {DDX_IPS_LAPLACE, "image_ps_laplace_"},
                {DDX_ISP, "image_sp_"},         {DDX_ISP_LAPLACE, "image_sp_laplace_"},
                {DDX_ISS, "image_ss_"},         {DDX_ISS_LAPLACE, "image_ss_laplace_"},
                {DDX_IP, "image_ip_"}};
    const int nout = arg_list.image_output ? 9 : 2;
    for (int q = 0; q < nout; q++) {
      check(ddx_get_image(ctx, outs[q].id, r, img.ptr()), ctx, "ddx_get_image");
      img.save(arg_list.output_path + outs[q].name + std::to_string(r), SaveFormat::binary_raw);
    }
  }
  ddx_destroy(ctx);
  return 0;
}
