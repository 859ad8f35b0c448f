// HIP kernels of the elastic FDTD / RTM engine, written for gfx950 (MI355X).
//
// All kernels are HBM-bound stencils or copies; none uses MFMA.  Every launch
// covers all subdomains (or all slabs / halo boxes) of a context at once
// through a descriptor table, so a time step costs a handful of launches
// regardless of the decomposition.  Blocks map to descriptors by a binary
// search over the descriptors' first-block indices.
//
// Arithmetic follows the reference's operation order (see each kernel's
// citation) and the library is compiled with -ffp-contract=off, so one FP32
// rounding happens per source-level operation, as in the CPU oracle.
#include "ddx_internal.h"

namespace ddx {

// Coefficients of half order h (rows 1..4), zero-padded to 4 taps.  Row 4 is
// phy_dc, rows R..4 are phy_dc_list (UTIL/.jarvis_2_geometry/
// jarvis_device_geometry_bones.hpp:29-71); row 2 is pml_dc.
__constant__ float c_fd[5][4];

void set_fd_table(const float tab[5][4]) {
  (void)hipMemcpyToSymbol(HIP_SYMBOL(c_fd), tab, sizeof(float) * 20);
}

__device__ __forceinline__ int find_desc(const int *__restrict__ starts, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (starts[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Forward-staggered: sum_r c[r] (f[i+(r+1)s] - f[i-r s]); backward-staggered:
// sum_r c[r] (f[i+r s] - f[i-(r+1)s]); accumulated from r = 0 upward
// (PHY/fdtd3d_phy_mpi_2_vel_kernel.hpp:26-50).
__device__ __forceinline__ float dp4(const float *__restrict__ f, long i, long s, float c0, float c1, float c2,
                                     float c3) {
  float a = 0.f;
  a += c0 * (f[i + s] - f[i]);
  a += c1 * (f[i + 2 * s] - f[i - s]);
  a += c2 * (f[i + 3 * s] - f[i - 2 * s]);
  a += c3 * (f[i + 4 * s] - f[i - 3 * s]);
  return a;
}
__device__ __forceinline__ float dm4(const float *__restrict__ f, long i, long s, float c0, float c1, float c2,
                                     float c3) {
  float a = 0.f;
  a += c0 * (f[i] - f[i - s]);
  a += c1 * (f[i + s] - f[i - 2 * s]);
  a += c2 * (f[i + 2 * s] - f[i - 3 * s]);
  a += c3 * (f[i + 3 * s] - f[i - 4 * s]);
  return a;
}
__device__ __forceinline__ float dp2(const float *__restrict__ f, long i, long s, float c0, float c1) {
  float a = 0.f;
  a += c0 * (f[i + s] - f[i]);
  a += c1 * (f[i + 2 * s] - f[i - s]);
  return a;
}
__device__ __forceinline__ float dm2(const float *__restrict__ f, long i, long s, float c0, float c1) {
  float a = 0.f;
  a += c0 * (f[i] - f[i - s]);
  a += c1 * (f[i + s] - f[i - 2 * s]);
  return a;
}

// VOWR variable order (PHY/fdtd3d_phy_mpi_2_vel_kernel.hpp:132-152, :240-260):
// in backward reconstruction a cell closer than phy_fdorder_half to the global
// physical boundary along an axis uses half order max(R, d+1) on that axis.
// Taps beyond the order carry zero coefficients; they read cells inside the
// extended margin (at least 4 cells wide), so the extra products are exact
// zeros and the sum equals the reference's shorter loop.
__device__ __forceinline__ int vowr_order(int p, int n, int R) {
  int d = min(p, n - 1 - p);
  return d >= 4 ? 4 : max(R, d + 1);
}

struct Coef4 {
  float c0, c1, c2, c3;
};
__device__ __forceinline__ Coef4 coef_row(int h) {
  Coef4 c;
  c.c0 = c_fd[h][0];
  c.c1 = c_fd[h][1];
  c.c2 = c_fd[h][2];
  c.c3 = c_fd[h][3];
  return c;
}

// ---------------------------------------------------------------------------
// Velocity update of the physical region
// (PHY/fdtd3d_phy_mpi_2_vel_kernel.hpp:7-80 inside; :82-2087 edges).
// MODE: PURE_FORWARD / RTM_FORWARD (no vxp), RTM_BACKWARD (variable order,
// vxp), RTM_REVERSE (vxp).
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_vel_phy(const PhyDesc *__restrict__ descs,
                                                    const int *__restrict__ starts, int n, float dx, float dy,
                                                    float dz, float sign, int gnx, int gny, int gnz, int R) {
  const int b = blockIdx.x;
  const PhyDesc &D = descs[find_desc(starts, n, b)];
  const int t = b - D.blk0;
  const int tiles = D.tiles_x * D.tiles_y;
  const int k = t / tiles;
  const int rem = t - k * tiles;
  const int ty = rem / D.tiles_x;
  const int tx = rem - ty * D.tiles_x;
  const int i = tx * kTileX + (threadIdx.x & (kTileX - 1));
  const int j = ty * kTileY + (threadIdx.x / kTileX);
  if (i >= D.nx || j >= D.ny) return;
  const long sy = D.sy, sz = D.sz;
  const long id = (long)(D.ox + i) + sy * (D.oy + j) + sz * (D.oz + k);

  Coef4 cx, cy, cz;
  if (MODE == RTM_BACKWARD) {
    cx = coef_row(vowr_order(D.gx + i, gnx, R));
    cy = coef_row(vowr_order(D.gy + j, gny, R));
    cz = coef_row(vowr_order(D.gz + k, gnz, R));
  } else {
    cx = coef_row(4);
    cy = cx;
    cz = cx;
  }
  const float *sxx = D.f[SXX], *syy = D.f[SYY], *szz = D.f[SZZ];
  const float *sxy = D.f[SXY], *sxz = D.f[SXZ], *syz = D.f[SYZ];

  const float dtr = sign * D.dtr[id];
  const float a = dtr / dx, bb = dtr / dy, c = dtr / dz;
  const float sxx_x = dp4(sxx, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
  const float syy_y = dp4(syy, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
  const float szz_z = dp4(szz, id, sz, cz.c0, cz.c1, cz.c2, cz.c3);
  const float syx_x = dm4(sxy, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
  const float sxy_y = dm4(sxy, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
  const float sxz_z = dm4(sxz, id, sz, cz.c0, cz.c1, cz.c2, cz.c3);
  const float szx_x = dm4(sxz, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
  const float szy_y = dm4(syz, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
  const float syz_z = dm4(syz, id, sz, cz.c0, cz.c1, cz.c2, cz.c3);
  D.f[VX][id] += a * sxx_x + bb * sxy_y + c * sxz_z;
  D.f[VY][id] += a * syx_x + bb * syy_y + c * syz_z;
  D.f[VZ][id] += a * szx_x + bb * szy_y + c * szz_z;
  if (MODE == RTM_BACKWARD || MODE == RTM_REVERSE) {
    const float *sau = D.f[SAU];
    const float sx = dp4(sau, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
    const float sy_ = dp4(sau, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
    const float sz_ = dp4(sau, id, sz, cz.c0, cz.c1, cz.c2, cz.c3);
    D.f[VXP][id] += a * sx;
    D.f[VYP][id] += bb * sy_;
    D.f[VZP][id] += c * sz_;
  }
}

// ---------------------------------------------------------------------------
// Stress update of the physical region
// (PHY/fdtd3d_phy_mpi_3_stress_kernel.hpp:7-60 inside; :62-1306 edges).
// sau is updated in the RTM modes only (:53-58).
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_stress_phy(const PhyDesc *__restrict__ descs,
                                                       const int *__restrict__ starts, int n, float dx,
                                                       float dy, float dz, float dt, int gnx, int gny, int gnz,
                                                       int R) {
  const int b = blockIdx.x;
  const PhyDesc &D = descs[find_desc(starts, n, b)];
  const int t = b - D.blk0;
  const int tiles = D.tiles_x * D.tiles_y;
  const int k = t / tiles;
  const int rem = t - k * tiles;
  const int ty = rem / D.tiles_x;
  const int tx = rem - ty * D.tiles_x;
  const int i = tx * kTileX + (threadIdx.x & (kTileX - 1));
  const int j = ty * kTileY + (threadIdx.x / kTileX);
  if (i >= D.nx || j >= D.ny) return;
  const long sy = D.sy, sz = D.sz;
  const long id = (long)(D.ox + i) + sy * (D.oy + j) + sz * (D.oz + k);

  Coef4 cx, cy, cz;
  if (MODE == RTM_BACKWARD) {
    cx = coef_row(vowr_order(D.gx + i, gnx, R));
    cy = coef_row(vowr_order(D.gy + j, gny, R));
    cz = coef_row(vowr_order(D.gz + k, gnz, R));
  } else {
    cx = coef_row(4);
    cy = cx;
    cz = cx;
  }
  const float *vx = D.f[VX], *vy = D.f[VY], *vz = D.f[VZ];
  const float a = dt / dx, bb = dt / dy, c = dt / dz;
  const float lm = D.lam[id], mu = D.mu[id];
  const float l2m = lm + 2.0f * mu;
  const float vx_x = dm4(vx, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
  const float vx_y = dp4(vx, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
  const float vx_z = dp4(vx, id, sz, cz.c0, cz.c1, cz.c2, cz.c3);
  const float vy_x = dp4(vy, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
  const float vy_y = dm4(vy, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
  const float vy_z = dp4(vy, id, sz, cz.c0, cz.c1, cz.c2, cz.c3);
  const float vz_x = dp4(vz, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
  const float vz_y = dp4(vz, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
  const float vz_z = dm4(vz, id, sz, cz.c0, cz.c1, cz.c2, cz.c3);
  D.f[SXX][id] += a * l2m * vx_x + bb * lm * vy_y + c * lm * vz_z;
  D.f[SYY][id] += a * lm * vx_x + bb * l2m * vy_y + c * lm * vz_z;
  D.f[SZZ][id] += a * lm * vx_x + bb * lm * vy_y + c * l2m * vz_z;
  D.f[SXY][id] += a * mu * vy_x + bb * mu * vx_y;
  D.f[SXZ][id] += a * mu * vz_x + c * mu * vx_z;
  D.f[SYZ][id] += bb * mu * vz_y + c * mu * vy_z;
  if (MODE != PURE_FORWARD) D.f[SAU][id] += a * l2m * vx_x + bb * l2m * vy_y + c * l2m * vz_z;
}

// ---------------------------------------------------------------------------
// CPML halves (CPML/fdtd3d_cpml_mpi_2_kernel_device_{x,y,z}_meat.hpp):
// 4th order; each derivative divided by the spacing; for each damped axis
// psi = b psi + a d and d += psi; then v += dt/rho * sum (velocity) or
// s += dt * (C : grad v) (stress).  Memory variables of an axis are touched
// only where the cell lies in that axis' damping range: elsewhere the
// reference's coefficient a is 0 and psi stays 0, so skipping is exact.
// psi layout per damped axis slot: vel (3) then stress (3).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void slab_cell(const SlabDesc &D, int b, long &c, long &ncell, int &ei, int &ej,
                                          int &ek) {
  ncell = (long)D.ni * D.nj * D.nk;
  c = (long)(b - D.blk0) * kBlock + threadIdx.x;
  long r = c / D.ni;
  ei = D.i0 + (int)(c - r * D.ni);
  ej = D.j0 + (int)(r % D.nj);
  ek = D.k0 + (int)(r / D.nj);
}

__global__ __launch_bounds__(kBlock) void k_vel_pml(const SlabDesc *__restrict__ descs,
                                                    const int *__restrict__ starts, int n, float dx, float dy,
                                                    float dz) {
  const int b = blockIdx.x;
  const SlabDesc &D = descs[find_desc(starts, n, b)];
  long c, ncell;
  int ei, ej, ek;
  slab_cell(D, b, c, ncell, ei, ej, ek);
  if (c >= ncell) return;
  const long sy = D.sy, sz = D.sz;
  const long id = (long)ei + sy * ej + sz * ek;
  const float c0 = c_fd[2][0], c1 = c_fd[2][1];
  const float *sxx = D.f[SXX], *syy = D.f[SYY], *szz = D.f[SZZ];
  const float *sxy = D.f[SXY], *sxz = D.f[SXZ], *syz = D.f[SYZ];
  const float dtr = D.dtr[id];
  float sxx_x = dp2(sxx, id, 1, c0, c1), syx_x = dm2(sxy, id, 1, c0, c1), szx_x = dm2(sxz, id, 1, c0, c1);
  float syy_y = dp2(syy, id, sy, c0, c1), sxy_y = dm2(sxy, id, sy, c0, c1), szy_y = dm2(syz, id, sy, c0, c1);
  float szz_z = dp2(szz, id, sz, c0, c1), sxz_z = dm2(sxz, id, sz, c0, c1), syz_z = dm2(syz, id, sz, c0, c1);
  sxx_x = sxx_x / dx; syy_y = syy_y / dy; szz_z = szz_z / dz;
  sxy_y = sxy_y / dy; sxz_z = sxz_z / dz; syx_x = syx_x / dx;
  syz_z = syz_z / dz; szx_x = szx_x / dx; szy_y = szy_y / dy;
  int slot = 0;
  if (D.axes & 1) {
    if (ei < D.xlo || ei >= D.xhi) {
      float *p = D.psi + (long)(slot * 6) * ncell + c;
      const float *q = D.coef + (long)(slot * 4) * ncell + c;
      const float a = q[0], bb = q[ncell], ah = q[2 * ncell], bh = q[3 * ncell];
      float p0 = bh * p[0] + ah * sxx_x;
      float p1 = bb * p[ncell] + a * syx_x;
      float p2 = bb * p[2 * ncell] + a * szx_x;
      p[0] = p0; p[ncell] = p1; p[2 * ncell] = p2;
      sxx_x += p0; syx_x += p1; szx_x += p2;
    }
    slot++;
  }
  if (D.axes & 2) {
    if (ej < D.ylo || ej >= D.yhi) {
      float *p = D.psi + (long)(slot * 6) * ncell + c;
      const float *q = D.coef + (long)(slot * 4) * ncell + c;
      const float a = q[0], bb = q[ncell], ah = q[2 * ncell], bh = q[3 * ncell];
      float p0 = bb * p[0] + a * sxy_y;
      float p1 = bh * p[ncell] + ah * syy_y;
      float p2 = bb * p[2 * ncell] + a * szy_y;
      p[0] = p0; p[ncell] = p1; p[2 * ncell] = p2;
      sxy_y += p0; syy_y += p1; szy_y += p2;
    }
    slot++;
  }
  if (D.axes & 4) {
    if (ek < D.zlo || ek >= D.zhi) {
      float *p = D.psi + (long)(slot * 6) * ncell + c;
      const float *q = D.coef + (long)(slot * 4) * ncell + c;
      const float a = q[0], bb = q[ncell], ah = q[2 * ncell], bh = q[3 * ncell];
      float p0 = bb * p[0] + a * sxz_z;
      float p1 = bb * p[ncell] + a * syz_z;
      float p2 = bh * p[2 * ncell] + ah * szz_z;
      p[0] = p0; p[ncell] = p1; p[2 * ncell] = p2;
      sxz_z += p0; syz_z += p1; szz_z += p2;
    }
  }
  D.f[VX][id] += dtr * (sxx_x + sxy_y + sxz_z);
  D.f[VY][id] += dtr * (syx_x + syy_y + syz_z);
  D.f[VZ][id] += dtr * (szx_x + szy_y + szz_z);
}

template <int WITH_SAU>
__global__ __launch_bounds__(kBlock) void k_stress_pml(const SlabDesc *__restrict__ descs,
                                                       const int *__restrict__ starts, int n, float dx,
                                                       float dy, float dz, float dt) {
  const int b = blockIdx.x;
  const SlabDesc &D = descs[find_desc(starts, n, b)];
  long c, ncell;
  int ei, ej, ek;
  slab_cell(D, b, c, ncell, ei, ej, ek);
  if (c >= ncell) return;
  const long sy = D.sy, sz = D.sz;
  const long id = (long)ei + sy * ej + sz * ek;
  const float c0 = c_fd[2][0], c1 = c_fd[2][1];
  const float *vx = D.f[VX], *vy = D.f[VY], *vz = D.f[VZ];
  float vx_x = dm2(vx, id, 1, c0, c1), vy_x = dp2(vy, id, 1, c0, c1), vz_x = dp2(vz, id, 1, c0, c1);
  float vy_y = dm2(vy, id, sy, c0, c1), vx_y = dp2(vx, id, sy, c0, c1), vz_y = dp2(vz, id, sy, c0, c1);
  float vz_z = dm2(vz, id, sz, c0, c1), vx_z = dp2(vx, id, sz, c0, c1), vy_z = dp2(vy, id, sz, c0, c1);
  const float lm = D.lam[id], mu = D.mu[id];
  const float l2m = lm + 2.0f * mu;
  vx_x = vx_x / dx; vy_y = vy_y / dy; vz_z = vz_z / dz;
  vy_x = vy_x / dx; vz_x = vz_x / dx; vx_y = vx_y / dy;
  vz_y = vz_y / dy; vx_z = vx_z / dz; vy_z = vy_z / dz;
  int slot = 0;
  if (D.axes & 1) {
    if (ei < D.xlo || ei >= D.xhi) {
      float *p = D.psi + (long)(slot * 6 + 3) * ncell + c;
      const float *q = D.coef + (long)(slot * 4) * ncell + c;
      const float a = q[0], bb = q[ncell], ah = q[2 * ncell], bh = q[3 * ncell];
      float p0 = bb * p[0] + a * vx_x;
      float p1 = bh * p[ncell] + ah * vy_x;
      float p2 = bh * p[2 * ncell] + ah * vz_x;
      p[0] = p0; p[ncell] = p1; p[2 * ncell] = p2;
      vx_x += p0; vy_x += p1; vz_x += p2;
    }
    slot++;
  }
  if (D.axes & 2) {
    if (ej < D.ylo || ej >= D.yhi) {
      float *p = D.psi + (long)(slot * 6 + 3) * ncell + c;
      const float *q = D.coef + (long)(slot * 4) * ncell + c;
      const float a = q[0], bb = q[ncell], ah = q[2 * ncell], bh = q[3 * ncell];
      float p0 = bh * p[0] + ah * vx_y;
      float p1 = bb * p[ncell] + a * vy_y;
      float p2 = bh * p[2 * ncell] + ah * vz_y;
      p[0] = p0; p[ncell] = p1; p[2 * ncell] = p2;
      vx_y += p0; vy_y += p1; vz_y += p2;
    }
    slot++;
  }
  if (D.axes & 4) {
    if (ek < D.zlo || ek >= D.zhi) {
      float *p = D.psi + (long)(slot * 6 + 3) * ncell + c;
      const float *q = D.coef + (long)(slot * 4) * ncell + c;
      const float a = q[0], bb = q[ncell], ah = q[2 * ncell], bh = q[3 * ncell];
      float p0 = bh * p[0] + ah * vx_z;
      float p1 = bh * p[ncell] + ah * vy_z;
      float p2 = bb * p[2 * ncell] + a * vz_z;
      p[0] = p0; p[ncell] = p1; p[2 * ncell] = p2;
      vx_z += p0; vy_z += p1; vz_z += p2;
    }
  }
  D.f[SXX][id] += dt * (l2m * vx_x + lm * vy_y + lm * vz_z);
  D.f[SYY][id] += dt * (lm * vx_x + l2m * vy_y + lm * vz_z);
  D.f[SZZ][id] += dt * (lm * vx_x + lm * vy_y + l2m * vz_z);
  if (WITH_SAU) D.f[SAU][id] += dt * (l2m * vx_x + l2m * vy_y + l2m * vz_z);
  D.f[SXY][id] += dt * mu * (vy_x + vx_y);
  D.f[SXZ][id] += dt * mu * (vz_x + vx_z);
  D.f[SYZ][id] += dt * mu * (vz_y + vy_z);
}

// ---------------------------------------------------------------------------
// Box <-> contiguous buffer copies: the CMM halo pack/unpack
// (OP/jarvis_2_mpicu_halo_comm_operator_meat.hpp:299-326 with
// FLD/jarvis_3_device_field_meat.hpp:23-60) and the VOWR boundary save /
// inject (BASE/fdtd3d_base_5_wave_recon_meat.hpp:89-133).  Boxes are walked
// x-fastest, so both sides of every copy are coalesced along x.
// ---------------------------------------------------------------------------
template <int DIR>
__global__ __launch_bounds__(kBlock) void k_copy(const CopyDesc *__restrict__ descs,
                                                 const int *__restrict__ starts, int n, float *__restrict__ buf) {
  const int b = blockIdx.x;
  const CopyDesc &D = descs[find_desc(starts, n, b)];
  const long cnt = (long)D.ni * D.nj * D.nk;
  const long e = (long)(b - D.blk0) * kBlock + threadIdx.x;
  if (e >= cnt) return;
  const long r = e / D.ni;
  const int i = (int)(e - r * D.ni);
  const int j = (int)(r % D.nj);
  const int k = (int)(r / D.nj);
  const long id = (long)(D.i0 + i) + D.sy * (D.j0 + j) + D.sz * (D.k0 + k);
  if (DIR > 0) buf[D.buf_off + e] = D.field[id];
  else D.field[id] = buf[D.buf_off + e];
}

// Explosive source (PHY/fdtd3d_phy_mpi_3_stress_kernel.hpp:1308-1364); one
// thread, launched only for the subdomain that owns the shot node.
__global__ void k_source(float *sxx, float *syy, float *szz, long off, const float *w, int it, float sign,
                         int src_mode) {
  const float v = w[it];
  if (sign > 0) {
    if (src_mode == 0) { sxx[off] += v; syy[off] += v; }
    szz[off] += v;
  } else {
    if (src_mode == 0) { sxx[off] -= v; syy[off] -= v; }
    szz[off] -= v;
  }
}

// Receiver sampling (BASE/fdtd3d_base_4_seisrecord_meat.hpp:57-77) and
// reverse-time injection (:79-107), grid-stride over all kept receivers of
// all subdomains (no 1024-receiver limit).  seis: [3][nrec][ntime].
__global__ void k_recv_sample(float *const *__restrict__ vptr, const int *__restrict__ rsub,
                              const long *__restrict__ roff, int nrec, int it, int ntime,
                              float *__restrict__ seis) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrec; r += gridDim.x * blockDim.x) {
    const int s = rsub[r];
    const long o = roff[r];
#pragma unroll
    for (int q = 0; q < 3; q++) seis[((long)q * nrec + r) * ntime + it] = vptr[s * 3 + q][o];
  }
}
__global__ void k_recv_inject(float *const *__restrict__ vptr, const int *__restrict__ rsub,
                              const long *__restrict__ roff, int nrec, int it, int ntime,
                              const float *__restrict__ seis) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrec; r += gridDim.x * blockDim.x) {
    const int s = rsub[r];
    const long o = roff[r];
#pragma unroll
    for (int q = 0; q < 3; q++) vptr[s * 3 + q][o] += seis[((long)q * nrec + r) * ntime + it];
  }
}

// ---------------------------------------------------------------------------
// Imaging: P/S-decomposed zero-lag cross-correlation
// (IMG/rtm3d_iamge_kernel.hpp:6-51), accumulated in T (float as the
// reference, or double).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void img_cell(const ImgDesc &D, int b, long &q, long &id, bool &ok) {
  const long np = (long)D.nx * D.ny * D.nz;
  q = (long)(b - D.blk0) * kBlock + threadIdx.x;
  ok = q < np;
  const long r = q / D.nx;
  const int i = (int)(q - r * D.nx);
  const int j = (int)(r % D.ny);
  const int k = (int)(r / D.ny);
  id = (long)(D.ox + i) + D.sy * (D.oy + j) + D.sz * (D.oz + k);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_image(const ImgDesc *__restrict__ descs,
                                                  const int *__restrict__ starts, int n) {
  const int b = blockIdx.x;
  const ImgDesc &D = descs[find_desc(starts, n, b)];
  long q, id;
  bool ok;
  img_cell(D, b, q, id, ok);
  if (!ok) return;
  const float pfx = D.b[3][id], pfy = D.b[4][id], pfz = D.b[5][id];
  const float sfx = D.b[0][id] - pfx, sfy = D.b[1][id] - pfy, sfz = D.b[2][id] - pfz;
  const float prx = D.r[3][id], pry = D.r[4][id], prz = D.r[5][id];
  const float srx = D.r[0][id] - prx, sry = D.r[1][id] - pry, srz = D.r[2][id] - prz;
  T *Ip = (T *)D.img[0], *Ipp = (T *)D.img[1], *Ips = (T *)D.img[2], *Isp = (T *)D.img[3], *Iss = (T *)D.img[4];
  if (sizeof(T) == sizeof(float)) {
    Ip[q] += pfx * pfx + pfy * pfy + pfz * pfz;
    Ipp[q] += pfx * prx + pfy * pry + pfz * prz;
    Ips[q] += pfx * srx + pfy * sry + pfz * srz;
    Isp[q] += sfx * prx + sfy * pry + sfz * prz;
    Iss[q] += sfx * srx + sfy * sry + sfz * srz;
  } else {
    const T PX = pfx, PY = pfy, PZ = pfz, SX = sfx, SY = sfy, SZ = sfz;
    const T px = prx, py = pry, pz = prz, sx = srx, sy = sry, sz = srz;
    Ip[q] += PX * PX + PY * PY + PZ * PZ;
    Ipp[q] += PX * px + PY * py + PZ * pz;
    Ips[q] += PX * sx + PY * sy + PZ * sz;
    Isp[q] += SX * px + SY * py + SZ * pz;
    Iss[q] += SX * sx + SY * sy + SZ * sz;
  }
}

// 7-point Laplacian of the uncompensated images with a 2-cell border left
// untouched (IMG/rtm3d_iamge_kernel.hpp:184-200), per subdomain gridphy.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_laplace(const ImgDesc *__restrict__ descs,
                                                    const int *__restrict__ starts, int n) {
  const int b = blockIdx.x;
  const ImgDesc &D = descs[find_desc(starts, n, b)];
  const long np = (long)D.nx * D.ny * D.nz;
  const long q = (long)(b - D.blk0) * kBlock + threadIdx.x;
  if (q >= np) return;
  const long r = q / D.nx;
  const int i = (int)(q - r * D.nx);
  const int j = (int)(r % D.ny);
  const int k = (int)(r / D.ny);
  if (i < 2 || i > D.nx - 3 || j < 2 || j > D.ny - 3 || k < 2 || k > D.nz - 3) return;
  const long sy = D.nx, sz = (long)D.nx * D.ny;
  for (int w = 0; w < 4; w++) {
    const T *I = (const T *)D.img[1 + w];
    T *L = (T *)D.img[5 + w];
    L[q] = I[q + 1] + I[q - 1] + I[q + sy] + I[q - sy] + I[q + sz] + I[q - sz] - T(6) * I[q];
  }
}

// Source-illumination compensation (IMG/rtm3d_iamge_kernel.hpp:118-135).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_compensate(const ImgDesc *__restrict__ descs,
                                                       const int *__restrict__ starts, int n) {
  const int b = blockIdx.x;
  const ImgDesc &D = descs[find_desc(starts, n, b)];
  const long np = (long)D.nx * D.ny * D.nz;
  const long q = (long)(b - D.blk0) * kBlock + threadIdx.x;
  if (q >= np) return;
  const T ip = ((const T *)D.img[0])[q];
  if ((double)ip > 3.4e-38) {
    for (int w = 0; w < 4; w++) {
      T *I = (T *)D.img[1 + w];
      I[q] = I[q] / ip;
    }
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
void launch_vel_phy(int mode, const PhyDesc *d, const int *starts, int n, int nblk, float dx, float dy,
                    float dz, float sign, int gnx, int gny, int gnz, int R, hipStream_t st) {
  if (nblk <= 0) return;
  dim3 g(nblk), bl(kBlock);
  switch (mode) {
    case PURE_FORWARD:
    case RTM_FORWARD:
      hipLaunchKernelGGL(k_vel_phy<PURE_FORWARD>, g, bl, 0, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R);
      break;
    case RTM_BACKWARD:
      hipLaunchKernelGGL(k_vel_phy<RTM_BACKWARD>, g, bl, 0, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R);
      break;
    default:
      hipLaunchKernelGGL(k_vel_phy<RTM_REVERSE>, g, bl, 0, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R);
  }
}

void launch_stress_phy(int mode, const PhyDesc *d, const int *starts, int n, int nblk, float dx, float dy,
                       float dz, float dt, int gnx, int gny, int gnz, int R, hipStream_t st) {
  if (nblk <= 0) return;
  dim3 g(nblk), bl(kBlock);
  switch (mode) {
    case PURE_FORWARD:
      hipLaunchKernelGGL(k_stress_phy<PURE_FORWARD>, g, bl, 0, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R);
      break;
    case RTM_FORWARD:
      hipLaunchKernelGGL(k_stress_phy<RTM_FORWARD>, g, bl, 0, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R);
      break;
    case RTM_BACKWARD:
      hipLaunchKernelGGL(k_stress_phy<RTM_BACKWARD>, g, bl, 0, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R);
      break;
    default:
      hipLaunchKernelGGL(k_stress_phy<RTM_REVERSE>, g, bl, 0, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R);
  }
}

void launch_vel_pml(const SlabDesc *d, const int *starts, int n, int nblk, float dx, float dy, float dz,
                    hipStream_t st) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_vel_pml, dim3(nblk), dim3(kBlock), 0, st, d, starts, n, dx, dy, dz);
}

void launch_stress_pml(int with_sau, const SlabDesc *d, const int *starts, int n, int nblk, float dx,
                       float dy, float dz, float dt, hipStream_t st) {
  if (nblk <= 0) return;
  if (with_sau)
    hipLaunchKernelGGL(k_stress_pml<1>, dim3(nblk), dim3(kBlock), 0, st, d, starts, n, dx, dy, dz, dt);
  else
    hipLaunchKernelGGL(k_stress_pml<0>, dim3(nblk), dim3(kBlock), 0, st, d, starts, n, dx, dy, dz, dt);
}

void launch_copy(int dir, const CopyDesc *d, const int *starts, int n, int nblk, float *buf, hipStream_t st) {
  if (nblk <= 0) return;
  if (dir > 0) hipLaunchKernelGGL(k_copy<1>, dim3(nblk), dim3(kBlock), 0, st, d, starts, n, buf);
  else hipLaunchKernelGGL(k_copy<-1>, dim3(nblk), dim3(kBlock), 0, st, d, starts, n, buf);
}

void launch_source(float *sxx, float *syy, float *szz, long off, const float *wavelet, int it, float sign,
                   int src_mode, hipStream_t st) {
  hipLaunchKernelGGL(k_source, dim3(1), dim3(1), 0, st, sxx, syy, szz, off, wavelet, it, sign, src_mode);
}

void launch_recv_sample(float *const *vptr, const int *rsub, const long *roff, int nrec, int it, int ntime,
                        float *seis, hipStream_t st) {
  if (nrec <= 0) return;
  int g = (nrec + kBlock - 1) / kBlock;
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(k_recv_sample, dim3(g), dim3(kBlock), 0, st, vptr, rsub, roff, nrec, it, ntime, seis);
}

void launch_recv_inject(float *const *vptr, const int *rsub, const long *roff, int nrec, int it, int ntime,
                        const float *seis, hipStream_t st) {
  if (nrec <= 0) return;
  int g = (nrec + kBlock - 1) / kBlock;
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(k_recv_inject, dim3(g), dim3(kBlock), 0, st, vptr, rsub, roff, nrec, it, ntime, seis);
}

void launch_image(int fp64, const ImgDesc *d, const int *starts, int n, int nblk, hipStream_t st) {
  if (nblk <= 0) return;
  if (fp64) hipLaunchKernelGGL(k_image<double>, dim3(nblk), dim3(kBlock), 0, st, d, starts, n);
  else hipLaunchKernelGGL(k_image<float>, dim3(nblk), dim3(kBlock), 0, st, d, starts, n);
}

void launch_image_post(int fp64, const ImgDesc *d, const int *starts, int n, int nblk, hipStream_t st) {
  if (nblk <= 0) return;
  if (fp64) {
    hipLaunchKernelGGL(k_laplace<double>, dim3(nblk), dim3(kBlock), 0, st, d, starts, n);
    hipLaunchKernelGGL(k_compensate<double>, dim3(nblk), dim3(kBlock), 0, st, d, starts, n);
  } else {
    hipLaunchKernelGGL(k_laplace<float>, dim3(nblk), dim3(kBlock), 0, st, d, starts, n);
    hipLaunchKernelGGL(k_compensate<float>, dim3(nblk), dim3(kBlock), 0, st, d, starts, n);
  }
}

}  // namespace ddx
