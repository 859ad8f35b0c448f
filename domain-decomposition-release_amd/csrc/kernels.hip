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

#include <algorithm>
#include <mutex>
#include <type_traits>

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

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with its own L2.  The remap hands every XCD one contiguous run of logical
// blocks (whole neighbouring tiles and planes), so stencil halos are re-read
// from the XCD's own L2 rather than from another die.
__device__ __forceinline__ int xcd_remap(int b, int nblk) {
  // Blocks are dealt to the 8 XCDs round-robin (block b runs on XCD b % 8).
  // XCD x receives per + (x < rem) blocks; give it the contiguous chunk
  // starting at x * per + min(x, rem), so neighbouring tiles share an L2.
  const int per = nblk >> 3, rem = nblk & 7;
  const int x = b & 7, idx = b >> 3;
  return x * per + min(x, rem) + idx;
}

// Group-interleaved variant: logical blocks form groups of g consecutive
// blocks, and the groups are dealt to the XCDs round-robin, so a block's
// stencil neighbours within the group share its L2 while every XCD gets an
// even share of a launch whose blocks differ in cost (the CPML slabs).
// Blocks past the last whole round of 8 groups keep their index; g <= 0 is
// the identity.  A bijection on [0, nblk).
__device__ __forceinline__ int xcd_group_remap(int b, int nblk, int g) {
  if (g <= 0) return b;
  const int full = nblk / (8 * g) * (8 * g);
  if (b >= full) return b;
  const int x = b & 7, i = b >> 3;
  return ((i / g) * 8 + x) * g + (i % g);
}

// Block order of the z-blocked bricks: 0 one contiguous chunk per XCD
// (xcd_remap), g > 0 groups of g blocks dealt round-robin (the engine sets
// 8: measured 1-3.5% faster than contiguous chunks on the 256x256x200 RTM
// shot, every group size from 1 to 16 within 1% of each other).
__constant__ int c_brick_group;
void set_brick_group(int g) { (void)hipMemcpyToSymbol(HIP_SYMBOL(c_brick_group), &g, sizeof(int)); }
__device__ __forceinline__ int brick_block(int b, int nblk) {
  return c_brick_group > 0 ? xcd_group_remap(b, nblk, c_brick_group) : xcd_remap(b, nblk);
}

// Field pointers reach the kernels through descriptor tables in global
// memory, so the compiler sees generic pointers and would emit FLAT loads
// with one 64-bit address computation per tap.  They always point into
// device global memory: casting them to address space 1 gives global loads,
// and lets contiguous x taps merge into dwordx2/x4 loads with immediate
// offsets.
typedef __attribute__((address_space(1))) float gfloat;
__device__ __forceinline__ gfloat *G(float *p) { return (gfloat *)p; }
__device__ __forceinline__ const gfloat *G(const float *p) { return (const gfloat *)p; }
__device__ __forceinline__ gfloat *G(void *p) { return (gfloat *)p; }
__device__ __forceinline__ const gfloat *G(const void *p) { return (const gfloat *)p; }

// ---------------------------------------------------------------------------
// Wavefield storage formats.  Fields, CPML memory variables and the VOWR
// store hold FP32 (F32), or 16 bits (F16: IEEE binary16, BF16: bfloat16) of
// the value times a power-of-two scale of its field class.  The stencils
// compute in FP32: loads widen exactly, stores round to nearest even.  The
// scales are folded into the per-kernel constants by the host (dt/rho, dt),
// so a kernel's FP32 operations are those of the FP32 path on scaled values.
// ---------------------------------------------------------------------------
struct F32 {};
struct F16 {};
struct BF16 {};
typedef __attribute__((address_space(1))) _Float16 ghalf;
typedef __attribute__((address_space(1))) unsigned short gbf16;
template <class S> struct GPtr;
template <> struct GPtr<F32> { typedef gfloat T; typedef float raw; };
template <> struct GPtr<F16> { typedef ghalf T; typedef _Float16 raw; };
template <> struct GPtr<BF16> { typedef gbf16 T; typedef unsigned short raw; };

__device__ __forceinline__ float widen(float v) { return v; }
__device__ __forceinline__ float widen(_Float16 v) { return (float)v; }
__device__ __forceinline__ float widen(unsigned short v) { return __uint_as_float((unsigned)v << 16); }
// FP32 -> bfloat16, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short bf16_rne(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float narrow(float x, const gfloat *) { return x; }
__device__ __forceinline__ _Float16 narrow(float x, const ghalf *) { return (_Float16)x; }
__device__ __forceinline__ unsigned short narrow(float x, const gbf16 *) { return bf16_rne(x); }

template <typename GT>
__device__ __forceinline__ float ld(const GT *p, long i) { return widen(p[i]); }
// store v; returns the value the field now holds
template <typename GT>
__device__ __forceinline__ float stq(GT *p, long i, float v) {
  const auto r = narrow(v, p);
  p[i] = r;
  return widen(r);
}
// the value a store of v into a field of p's format would leave (exact
// round trip for FP32)
template <typename GT>
__device__ __forceinline__ float quant(const GT *p, float v) {
  return widen(narrow(v, p));
}
template <typename GT>
__device__ __forceinline__ void add_to(GT *p, long i, float v) {
  p[i] = narrow(widen(p[i]) + v, p);
}

// Cells that one thread alone reads and writes in a launch (the updated
// fields and the per-cell model arrays of the physical-region bricks, the
// imaging inputs and images) go through non-temporal loads and stores, so
// they do not push the stencil halos of neighbouring bricks out of L2.  Same
// arithmetic as `p[i] += v`.  (Measured: 2-4% faster bricks, imaging
// 259 -> 174 us per launch; slower in the CPML kernels and the halo copy,
// which keep ordinary accesses.)
template <typename GT>
__device__ __forceinline__ void rmw_add(GT *p, long i, float v) {
  const float old = widen(__builtin_nontemporal_load(p + i));
  __builtin_nontemporal_store(narrow(old + v, p), p + i);
}
template <typename GT>
__device__ __forceinline__ float ld_stream(const GT *p, long i) { return widen(__builtin_nontemporal_load(p + i)); }
template <typename GT>
__device__ __forceinline__ void st_stream(GT *p, long i, float v) { __builtin_nontemporal_store(narrow(v, p), p + i); }
template <typename GT, typename T>
__device__ __forceinline__ void rmw_add_t(GT *p, long i, T v) {
  const T old = __builtin_nontemporal_load(p + i);
  __builtin_nontemporal_store(old + v, p + i);
}

// ---------------------------------------------------------------------------
// Pack on store (MirrorDesc): after a kernel stores an exchanged field at a
// cell with local extended coordinates e, the same value goes into the
// ghost layers of every receiver that takes the cell.  Subdomains are at
// least 8 cells wide and halos at most 4, so a cell can be sent across at
// most one side per axis: mir_cell finds that candidate side per axis once
// per cell, and mir_put checks the field's width and stores.
// ---------------------------------------------------------------------------
struct MirCell {
  const MirrorDesc *M;
  int e[3];
  int sd[3];  // candidate side per axis (x: 4 lo / 5 hi, y: 2 / 3, z: 0 / 1), -1: none
  bool any;
};
__device__ __forceinline__ MirCell mir_cell(const MirrorDesc *__restrict__ M, int ex, int ey, int ez) {
  MirCell m;
  m.M = M;
  m.e[0] = ex; m.e[1] = ey; m.e[2] = ez;
  m.any = false;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int lo = 4 - 2 * a, c = m.e[a];
    int sd = -1;
    if (c < 4 && M->s[lo].base) sd = lo;
    else if (c >= M->en[a] - 4 && M->s[lo + 1].base) sd = lo + 1;
    m.sd[a] = sd;
    m.any |= sd >= 0;
  }
  return m;
}
template <typename GT>
__device__ __forceinline__ void mir_put(const MirCell &m, int f, float v) {
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const int sd = m.sd[a];
    if (sd < 0) continue;
    const MirrorSide &s = m.M->s[sd];
    const int c = m.e[a], w = s.w[f];
    if ((sd & 1) ? (c < m.M->en[a] - w) : (c >= w)) continue;
    const int r0 = m.e[0] + (a == 0 ? s.shift : 0);
    const int r1 = m.e[1] + (a == 1 ? s.shift : 0);
    const int r2 = m.e[2] + (a == 2 ? s.shift : 0);
    GT *p = (GT *)(void *)s.base;
    stq(p, (long)f * s.fpitch + r0 + s.sy * r1 + s.sz * r2, v);
  }
}
// the two cells of a pair (the second only when it is in gridphy)
template <typename GT>
__device__ __forceinline__ void mir_pair(const MirCell &ma, const MirCell &mb, bool both, int f, float va,
                                         float vb) {
  if (ma.any) mir_put<GT>(ma, f, va);
  if (both && mb.any) mir_put<GT>(mb, f, vb);
}


// Forward-staggered: sum_r c[r] (f[i+(r+1)s] - f[i-r s]); backward-staggered:
// sum_r c[r] (f[i+r s] - f[i-(r+1)s]); accumulated from r = 0 upward
// (PHY/fdtd3d_phy_mpi_2_vel_kernel.hpp:26-50).
template <typename PT>
__device__ __forceinline__ float dp4(PT f, long i, long s, float c0, float c1, float c2, float c3) {
  float a = 0.f;
  a += c0 * (ld(f, i + s) - ld(f, i));
  a += c1 * (ld(f, i + 2 * s) - ld(f, i - s));
  a += c2 * (ld(f, i + 3 * s) - ld(f, i - 2 * s));
  a += c3 * (ld(f, i + 4 * s) - ld(f, i - 3 * s));
  return a;
}
template <typename PT>
__device__ __forceinline__ float dm4(PT f, long i, long s, float c0, float c1, float c2, float c3) {
  float a = 0.f;
  a += c0 * (ld(f, i) - ld(f, i - s));
  a += c1 * (ld(f, i + s) - ld(f, i - 2 * s));
  a += c2 * (ld(f, i + 2 * s) - ld(f, i - 3 * s));
  a += c3 * (ld(f, i + 3 * s) - ld(f, i - 4 * s));
  return a;
}
template <typename PT>
__device__ __forceinline__ float dp2(PT f, long i, long s, float c0, float c1) {
  float a = 0.f;
  a += c0 * (ld(f, i + s) - ld(f, i));
  a += c1 * (ld(f, i + 2 * s) - ld(f, i - s));
  return a;
}
template <typename PT>
__device__ __forceinline__ float dm2(PT f, long i, long s, float c0, float c1) {
  float a = 0.f;
  a += c0 * (ld(f, i) - ld(f, i - s));
  a += c1 * (ld(f, i + s) - ld(f, i - 2 * s));
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
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const PhyDesc &D = descs[find_desc(starts, n, b)];
  gnx = D.gnx; gny = D.gny; gnz = D.gnz; R = D.R;  // per descriptor: fused batches mix grids
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
  const gfloat *sxx = G(D.f[SXX]), *syy = G(D.f[SYY]), *szz = G(D.f[SZZ]);
  const gfloat *sxy = G(D.f[SXY]), *sxz = G(D.f[SXZ]), *syz = G(D.f[SYZ]);

  const float dtr = sign * G(D.dtr)[id];
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
  G(D.f[VX])[id] += a * sxx_x + bb * sxy_y + c * sxz_z;
  G(D.f[VY])[id] += a * syx_x + bb * syy_y + c * syz_z;
  G(D.f[VZ])[id] += a * szx_x + bb * szy_y + c * szz_z;
  if (MODE == RTM_BACKWARD || MODE == RTM_REVERSE) {
    const gfloat *sau = G(D.f[SAU]);
    const float sx = dp4(sau, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
    const float sy_ = dp4(sau, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
    const float sz_ = dp4(sau, id, sz, cz.c0, cz.c1, cz.c2, cz.c3);
    G(D.f[VXP])[id] += a * sx;
    G(D.f[VYP])[id] += bb * sy_;
    G(D.f[VZP])[id] += c * sz_;
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
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const PhyDesc &D = descs[find_desc(starts, n, b)];
  gnx = D.gnx; gny = D.gny; gnz = D.gnz; R = D.R;  // per descriptor: fused batches mix grids
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
  const gfloat *vx = G(D.f[VX]), *vy = G(D.f[VY]), *vz = G(D.f[VZ]);
  const float a = dt / dx, bb = dt / dy, c = dt / dz;
  const float lm = G(D.lam)[id], mu = G(D.mu)[id];
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
  G(D.f[SXX])[id] += a * l2m * vx_x + bb * lm * vy_y + c * lm * vz_z;
  G(D.f[SYY])[id] += a * lm * vx_x + bb * l2m * vy_y + c * lm * vz_z;
  G(D.f[SZZ])[id] += a * lm * vx_x + bb * lm * vy_y + c * l2m * vz_z;
  G(D.f[SXY])[id] += a * mu * vy_x + bb * mu * vx_y;
  G(D.f[SXZ])[id] += a * mu * vz_x + c * mu * vx_z;
  G(D.f[SYZ])[id] += bb * mu * vz_y + c * mu * vy_z;
  if (rtm_mode(MODE)) G(D.f[SAU])[id] += a * l2m * vx_x + bb * l2m * vy_y + c * l2m * vz_z;
}

// ---------------------------------------------------------------------------
// z-blocked versions of the physical-region updates.  A thread computes CZ
// consecutive planes of CY columns; the z arms of the stencil come from a
// register column of CZ+7 planes loaded once (instead of 8 loads per output),
// and every load of the block is independent of the others, so the whole
// column is in flight at once.  x and y arms are read through the caches.
// Blocks own TX x (TY*CY) x CZ bricks; the XCD remap keeps neighbouring
// bricks on one L2.  Operands and order of operations are those of
// k_vel_phy / k_stress_phy, so results are bit-identical.
// ---------------------------------------------------------------------------
template <int TX, int TY, int CY, int CZ>
__device__ __forceinline__ bool zbrick(const PhyDesc &D, int t, int &i, int &j0, int &k0) {
  const int tiles_x = (D.nx + TX - 1) / TX;
  const int tiles_y = (D.ny + TY * CY - 1) / (TY * CY);
  const int per = tiles_x * tiles_y;
  const int zc = t / per;
  const int rem = t - zc * per;
  const int ty = rem / tiles_x;
  const int tx = rem - ty * tiles_x;
  i = tx * TX + (int)(threadIdx.x % TX);
  j0 = ty * TY * CY + (int)(threadIdx.x / TX);
  k0 = zc * CZ;
  return i < D.nx;
}

// zcol[r] = f(plane k0 + lo + r), r < CZ + 7, planes past nz + 3 read as 0
template <int CZ, typename GT>
__device__ __forceinline__ void load_zcol(float (&zc)[CZ + 7], const GT *__restrict__ f, long base, long sz,
                                          int k0, int lo, int nz) {
#pragma unroll
  for (int r = 0; r < CZ + 7; r++) zc[r] = (k0 + lo + r < nz + 4) ? ld(f, base + (long)(lo + r) * sz) : 0.f;
}

// x arm from neighbouring lanes (XL bricks, tx = 64: a wavefront is one x
// row of 64 cells).  Each lane holds its own cell c; the taps c(x +- d),
// d = 1..4, arrive by whole-wavefront one-lane shifts (DPP wave_shr /
// wave_shl), and the lanes that shift off the row edge take the halo
// register g instead (lanes 60..63: cells x0-4 .. x0-1; lanes 0..3: cells
// x0+64 .. x0+67), rotated into place (wave_ror / wave_rol) one step ahead.
// Exact moves: the taps are the loaded values bit for bit.
enum { DPP_SHL1 = 0x130, DPP_ROL1 = 0x134, DPP_SHR1 = 0x138, DPP_ROR1 = 0x13C };
template <int CTL>
__device__ __forceinline__ float dpp_mov(float old, float src) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTL, 0xF, 0xF, false));
}
struct XRow {
  float m4, m3, m2, m1, c, p1, p2, p3, p4;
};
__device__ __forceinline__ XRow xrow(float c, float g) {
  XRow r;
  r.c = c;
  float gl = dpp_mov<DPP_ROR1>(0.f, g), gr = dpp_mov<DPP_ROL1>(0.f, g);
  r.m1 = dpp_mov<DPP_SHR1>(gl, c);
  r.p1 = dpp_mov<DPP_SHL1>(gr, c);
  gl = dpp_mov<DPP_ROR1>(0.f, gl);
  gr = dpp_mov<DPP_ROL1>(0.f, gr);
  r.m2 = dpp_mov<DPP_SHR1>(gl, r.m1);
  r.p2 = dpp_mov<DPP_SHL1>(gr, r.p1);
  gl = dpp_mov<DPP_ROR1>(0.f, gl);
  gr = dpp_mov<DPP_ROL1>(0.f, gr);
  r.m3 = dpp_mov<DPP_SHR1>(gl, r.m2);
  r.p3 = dpp_mov<DPP_SHL1>(gr, r.p2);
  gl = dpp_mov<DPP_ROR1>(0.f, gl);
  gr = dpp_mov<DPP_ROL1>(0.f, gr);
  r.m4 = dpp_mov<DPP_SHR1>(gl, r.m3);
  r.p4 = dpp_mov<DPP_SHL1>(gr, r.p3);
  return r;
}
// dp4 / dm4 along x with the same operands and order
__device__ __forceinline__ float dpx(const XRow &r, const Coef4 &c) {
  float a = 0.f;
  a += c.c0 * (r.p1 - r.c);
  a += c.c1 * (r.p2 - r.m1);
  a += c.c2 * (r.p3 - r.m2);
  a += c.c3 * (r.p4 - r.m3);
  return a;
}
__device__ __forceinline__ float dmx(const XRow &r, const Coef4 &c) {
  float a = 0.f;
  a += c.c0 * (r.c - r.m1);
  a += c.c1 * (r.p1 - r.m2);
  a += c.c2 * (r.p2 - r.m3);
  a += c.c3 * (r.p3 - r.m4);
  return a;
}
// Row geometry of an XL brick lane: the cell it loads (clamped to the
// margin: lanes past gridphy + 3 read a valid cell and store nothing) and
// its halo cell.
struct XLane {
  int ic, hx;
  bool halo;
};
__device__ __forceinline__ XLane xlane(int i, int nx) {
  const int lane = i & 63, x0 = i - lane;
  XLane r;
  r.ic = min(i, nx + 3);
  r.halo = lane < 4 || lane >= 60;
  r.hx = lane < 4 ? min(x0 + 64 + lane, nx + 3) : x0 + lane - 64;
  return r;
}
template <typename GT>
__device__ __forceinline__ float ld_halo(const GT *p, const XLane &xl, long id_h) {
  float g = 0.f;
  if (xl.halo) g = ld(p, id_h);
  return g;
}

// S: storage format of the fields; `sign` carries the velocity-over-stress
// scale ratio of the 16-bit formats (a power of two, exact).  XL: x arm from
// neighbouring lanes (tx = 64); `live`: the lane's cell is in gridphy.
template <int MODE, int TX, int TY, int CY, int CZ, class S, bool XL = false>
__device__ __forceinline__ void vel_zb_body(const PhyDesc &D, int i, int j0, int k0, float dx, float dy, float dz,
                                            float sign, int gnx, int gny, int gnz, int R, bool live = true) {
  typedef typename GPtr<S>::T gT;
  const long sy = D.sy, sz = D.sz;
  const XLane xl = xlane(i, D.nx);
  const int ic = XL ? xl.ic : i;
  const gT *__restrict__ sxx = (const gT *)D.f[SXX];
  const gT *__restrict__ syy = (const gT *)D.f[SYY];
  const gT *__restrict__ szz = (const gT *)D.f[SZZ];
  const gT *__restrict__ sxy = (const gT *)D.f[SXY];
  const gT *__restrict__ sxz = (const gT *)D.f[SXZ];
  const gT *__restrict__ syz = (const gT *)D.f[SYZ];
  const gT *__restrict__ sau = (const gT *)D.f[SAU];
  gT *__restrict__ fvx = (gT *)D.f[VX];
  gT *__restrict__ fvy = (gT *)D.f[VY];
  gT *__restrict__ fvz = (gT *)D.f[VZ];
  gT *__restrict__ fvxp = (gT *)D.f[VXP];
  gT *__restrict__ fvyp = (gT *)D.f[VYP];
  gT *__restrict__ fvzp = (gT *)D.f[VZP];
  const gfloat *__restrict__ fdtr = G(D.dtr);
  constexpr bool P = (MODE == RTM_BACKWARD || MODE == RTM_REVERSE);
  const Coef4 c4 = coef_row(4);
  const Coef4 cx = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gx + i, gnx, R)) : c4;
#pragma unroll
  for (int q = 0; q < CY; q++) {
    const int j = j0 + q * TY;
    if (j >= D.ny) continue;
    const Coef4 cy = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gy + j, gny, R)) : c4;
    const long base = (long)(D.ox + ic) + sy * (D.oy + j) + sz * (D.oz + k0);
    const long hbase = base + (xl.hx - ic);
    float zzz[CZ + 7], zxz[CZ + 7], zyz[CZ + 7], zsa[P ? CZ + 7 : 1];
    load_zcol<CZ>(zzz, szz, base, sz, k0, -3, D.nz);  // planes k0-3 ..
    load_zcol<CZ>(zxz, sxz, base, sz, k0, -4, D.nz);  // planes k0-4 ..
    load_zcol<CZ>(zyz, syz, base, sz, k0, -4, D.nz);
    if constexpr (P) load_zcol<CZ>(zsa, sau, base, sz, k0, -3, D.nz);
#pragma unroll
    for (int kk = 0; kk < CZ; kk++) {
      const int k = k0 + kk;
      if (k >= D.nz) break;
      const Coef4 cz = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gz + k, gnz, R)) : c4;
      const long id = base + (long)kk * sz;
      // the updated cells are read before the stencil taps (see stress_zb_body)
      const float ovx = ld_stream(fvx, id), ovy = ld_stream(fvy, id), ovz = ld_stream(fvz, id);
      const float ovxp = P ? ld_stream(fvxp, id) : 0.f, ovyp = P ? ld_stream(fvyp, id) : 0.f;
      const float ovzp = P ? ld_stream(fvzp, id) : 0.f;
      const float dtr = sign * ld_stream(fdtr, id);
      const float a = dtr / dx, bb = dtr / dy, c = dtr / dz;
      // XL: each x row is expanded and reduced to its derivative at once,
      // so one row (9 values) is live at a time
      float xl_xx = 0.f, xl_yx = 0.f, xl_zx = 0.f, xl_sa = 0.f;
      if constexpr (XL) {
        const long hid = hbase + (long)kk * sz;
        const float gxx = ld_halo(sxx, xl, hid), gxy = ld_halo(sxy, xl, hid), gxz = ld_halo(sxz, xl, hid);
        const float gsa = P ? ld_halo(sau, xl, hid) : 0.f;
        const float cxx = ld(sxx, id), cxy = ld(sxy, id);
        xl_xx = dpx(xrow(cxx, gxx), cx);
        xl_yx = dmx(xrow(cxy, gxy), cx);
        xl_zx = dmx(xrow(zxz[kk + 4], gxz), cx);
        if constexpr (P) xl_sa = dpx(xrow(zsa[kk + 3], gsa), cx);
      }
      const float sxx_x = XL ? xl_xx : dp4(sxx, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
      const float syy_y = dp4(syy, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
      // forward-staggered z arm: f(k+r+1) - f(k-r) -> zzz[kk+4+r] - zzz[kk+3-r]
      float szz_z = 0.f;
      szz_z += cz.c0 * (zzz[kk + 4] - zzz[kk + 3]);
      szz_z += cz.c1 * (zzz[kk + 5] - zzz[kk + 2]);
      szz_z += cz.c2 * (zzz[kk + 6] - zzz[kk + 1]);
      szz_z += cz.c3 * (zzz[kk + 7] - zzz[kk + 0]);
      const float syx_x = XL ? xl_yx : dm4(sxy, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
      const float sxy_y = dm4(sxy, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
      // backward-staggered z arm: f(k+r) - f(k-r-1) -> zxz[kk+4+r] - zxz[kk+3-r]
      float sxz_z = 0.f;
      sxz_z += cz.c0 * (zxz[kk + 4] - zxz[kk + 3]);
      sxz_z += cz.c1 * (zxz[kk + 5] - zxz[kk + 2]);
      sxz_z += cz.c2 * (zxz[kk + 6] - zxz[kk + 1]);
      sxz_z += cz.c3 * (zxz[kk + 7] - zxz[kk + 0]);
      const float szx_x = XL ? xl_zx : dm4(sxz, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
      const float szy_y = dm4(syz, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
      float syz_z = 0.f;
      syz_z += cz.c0 * (zyz[kk + 4] - zyz[kk + 3]);
      syz_z += cz.c1 * (zyz[kk + 5] - zyz[kk + 2]);
      syz_z += cz.c2 * (zyz[kk + 6] - zyz[kk + 1]);
      syz_z += cz.c3 * (zyz[kk + 7] - zyz[kk + 0]);
      float sx = 0.f, sy_ = 0.f, sz_ = 0.f;
      if constexpr (P) {
        sx = XL ? xl_sa : dp4(sau, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
        sy_ = dp4(sau, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
        sz_ += cz.c0 * (zsa[kk + 4] - zsa[kk + 3]);
        sz_ += cz.c1 * (zsa[kk + 5] - zsa[kk + 2]);
        sz_ += cz.c2 * (zsa[kk + 6] - zsa[kk + 1]);
        sz_ += cz.c3 * (zsa[kk + 7] - zsa[kk + 0]);
      }
      if (XL && !live) continue;
      if constexpr (MODE == FWI) {
        // PHY/fdtd3d_phy_mpi_2_vel_kernel.hpp:70-78: u += dt v (after the
        // update), vrt = dt x the increment
        const float dt = D.dt;
        const float oux = ld_stream(G(D.f[UX]), id), ouy = ld_stream(G(D.f[UY]), id), ouz = ld_stream(G(D.f[UZ]), id);
        const float ix = a * sxx_x + bb * sxy_y + c * sxz_z;
        const float iy = a * syx_x + bb * syy_y + c * syz_z;
        const float iz = a * szx_x + bb * szy_y + c * szz_z;
        const float nvx = ovx + ix, nvy = ovy + iy, nvz = ovz + iz;
        st_stream(fvx, id, nvx);
        st_stream(fvy, id, nvy);
        st_stream(fvz, id, nvz);
        st_stream(G(D.f[UX]), id, oux + dt * nvx);
        st_stream(G(D.f[UY]), id, ouy + dt * nvy);
        st_stream(G(D.f[UZ]), id, ouz + dt * nvz);
        __builtin_nontemporal_store(dt * ix, G(D.f[VXRT]) + id);
        __builtin_nontemporal_store(dt * iy, G(D.f[VYRT]) + id);
        __builtin_nontemporal_store(dt * iz, G(D.f[VZRT]) + id);
        if (D.mir) {
          const MirCell m = mir_cell(D.mir, D.ox + i, D.oy + j, D.oz + k);
          if (m.any) { mir_put<gT>(m, VX, nvx); mir_put<gT>(m, VY, nvy); mir_put<gT>(m, VZ, nvz); }
        }
      } else {
        const float nvx = ovx + (a * sxx_x + bb * sxy_y + c * sxz_z);
        const float nvy = ovy + (a * syx_x + bb * syy_y + c * syz_z);
        const float nvz = ovz + (a * szx_x + bb * szy_y + c * szz_z);
        st_stream(fvx, id, nvx);
        st_stream(fvy, id, nvy);
        st_stream(fvz, id, nvz);
        if (D.mir) {
          const MirCell m = mir_cell(D.mir, D.ox + i, D.oy + j, D.oz + k);
          if (m.any) { mir_put<gT>(m, VX, nvx); mir_put<gT>(m, VY, nvy); mir_put<gT>(m, VZ, nvz); }
        }
      }
      if constexpr (P) {
        st_stream(fvxp, id, ovxp + a * sx);
        st_stream(fvyp, id, ovyp + bb * sy_);
        st_stream(fvzp, id, ovzp + c * sz_);
      }
    }
  }
}

// `dt` carries the stress-over-velocity scale ratio of the 16-bit formats.
template <int MODE, int TX, int TY, int CY, int CZ, class S, bool XL = false>
__device__ __forceinline__ void stress_zb_body(const PhyDesc &D, int i, int j0, int k0, float dx, float dy,
                                               float dz, float dt, int gnx, int gny, int gnz, int R,
                                               bool live = true) {
  typedef typename GPtr<S>::T gT;
  const long sy = D.sy, sz = D.sz;
  const XLane xl = xlane(i, D.nx);
  const int ic = XL ? xl.ic : i;
  const gT *__restrict__ vx = (const gT *)D.f[VX];
  const gT *__restrict__ vy = (const gT *)D.f[VY];
  const gT *__restrict__ vz = (const gT *)D.f[VZ];
  gT *__restrict__ fsxx = (gT *)D.f[SXX];
  gT *__restrict__ fsyy = (gT *)D.f[SYY];
  gT *__restrict__ fszz = (gT *)D.f[SZZ];
  gT *__restrict__ fsxy = (gT *)D.f[SXY];
  gT *__restrict__ fsxz = (gT *)D.f[SXZ];
  gT *__restrict__ fsyz = (gT *)D.f[SYZ];
  gT *__restrict__ fsau = (gT *)D.f[SAU];
  const gfloat *__restrict__ flam = G(D.lam);
  const gfloat *__restrict__ fmu = G(D.mu);
  const float a = dt / dx, bb = dt / dy, c = dt / dz;
  const Coef4 c4 = coef_row(4);
  const Coef4 cx = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gx + i, gnx, R)) : c4;
#pragma unroll
  for (int q = 0; q < CY; q++) {
    const int j = j0 + q * TY;
    if (j >= D.ny) continue;
    const Coef4 cy = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gy + j, gny, R)) : c4;
    const long base = (long)(D.ox + ic) + sy * (D.oy + j) + sz * (D.oz + k0);
    const long hbase = base + (xl.hx - ic);
    float zx[CZ + 7], zy[CZ + 7], zz[CZ + 7];
    load_zcol<CZ>(zx, vx, base, sz, k0, -3, D.nz);  // vx, vy: forward-staggered in z
    load_zcol<CZ>(zy, vy, base, sz, k0, -3, D.nz);
    load_zcol<CZ>(zz, vz, base, sz, k0, -4, D.nz);  // vz: backward-staggered in z
#pragma unroll
    for (int kk = 0; kk < CZ; kk++) {
      const int k = k0 + kk;
      if (k >= D.nz) break;
      const Coef4 cz = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gz + k, gnz, R)) : c4;
      const long id = base + (long)kk * sz;
      // the updated cells are read before the stencil taps: on this target
      // the memory counter also tracks stores, so a read issued after a
      // store waits for that store, which would serialise the updates
      const float oxx = ld_stream(fsxx, id), oyy = ld_stream(fsyy, id), ozz = ld_stream(fszz, id);
      const float oxy = ld_stream(fsxy, id), oxz = ld_stream(fsxz, id), oyz = ld_stream(fsyz, id);
      const float oau = rtm_mode(MODE) ? ld_stream(fsau, id) : 0.f;
      const float lm = ld_stream(flam, id), mu = ld_stream(fmu, id);
      const float l2m = lm + 2.0f * mu;
      float xl_x = 0.f, xl_y = 0.f, xl_z = 0.f;  // x derivatives of vx, vy, vz (XL)
      if constexpr (XL) {
        const long hid = hbase + (long)kk * sz;
        const float gx = ld_halo(vx, xl, hid), gy = ld_halo(vy, xl, hid), gz = ld_halo(vz, xl, hid);
        xl_x = dmx(xrow(zx[kk + 3], gx), cx);
        xl_y = dpx(xrow(zy[kk + 3], gy), cx);
        xl_z = dpx(xrow(zz[kk + 4], gz), cx);
      }
      const float vx_x = XL ? xl_x : dm4(vx, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
      const float vx_y = dp4(vx, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
      float vx_z = 0.f;
      vx_z += cz.c0 * (zx[kk + 4] - zx[kk + 3]);
      vx_z += cz.c1 * (zx[kk + 5] - zx[kk + 2]);
      vx_z += cz.c2 * (zx[kk + 6] - zx[kk + 1]);
      vx_z += cz.c3 * (zx[kk + 7] - zx[kk + 0]);
      const float vy_x = XL ? xl_y : dp4(vy, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
      const float vy_y = dm4(vy, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
      float vy_z = 0.f;
      vy_z += cz.c0 * (zy[kk + 4] - zy[kk + 3]);
      vy_z += cz.c1 * (zy[kk + 5] - zy[kk + 2]);
      vy_z += cz.c2 * (zy[kk + 6] - zy[kk + 1]);
      vy_z += cz.c3 * (zy[kk + 7] - zy[kk + 0]);
      const float vz_x = XL ? xl_z : dp4(vz, id, 1, cx.c0, cx.c1, cx.c2, cx.c3);
      const float vz_y = dp4(vz, id, sy, cy.c0, cy.c1, cy.c2, cy.c3);
      float vz_z = 0.f;
      vz_z += cz.c0 * (zz[kk + 4] - zz[kk + 3]);
      vz_z += cz.c1 * (zz[kk + 5] - zz[kk + 2]);
      vz_z += cz.c2 * (zz[kk + 6] - zz[kk + 1]);
      vz_z += cz.c3 * (zz[kk + 7] - zz[kk + 0]);
      if (XL && !live) continue;
      const float nxx = oxx + (a * l2m * vx_x + bb * lm * vy_y + c * lm * vz_z);
      const float nyy = oyy + (a * lm * vx_x + bb * l2m * vy_y + c * lm * vz_z);
      const float nzz = ozz + (a * lm * vx_x + bb * lm * vy_y + c * l2m * vz_z);
      const float nxy = oxy + (a * mu * vy_x + bb * mu * vx_y);
      const float nxz = oxz + (a * mu * vz_x + c * mu * vx_z);
      const float nyz = oyz + (bb * mu * vz_y + c * mu * vy_z);
      const float nau = rtm_mode(MODE) ? oau + (a * l2m * vx_x + bb * l2m * vy_y + c * l2m * vz_z) : 0.f;
      st_stream(fsxx, id, nxx);
      st_stream(fsyy, id, nyy);
      st_stream(fszz, id, nzz);
      st_stream(fsxy, id, nxy);
      st_stream(fsxz, id, nxz);
      st_stream(fsyz, id, nyz);
      if constexpr (rtm_mode(MODE)) st_stream(fsau, id, nau);
      if (D.mir) {
        const MirCell m = mir_cell(D.mir, D.ox + i, D.oy + j, D.oz + k);
        if (m.any) {
          mir_put<gT>(m, SXX, nxx); mir_put<gT>(m, SYY, nyy); mir_put<gT>(m, SZZ, nzz);
          mir_put<gT>(m, SXY, nxy); mir_put<gT>(m, SXZ, nxz); mir_put<gT>(m, SYZ, nyz);
          if constexpr (rtm_mode(MODE)) mir_put<gT>(m, SAU, nau);
        }
      }
    }
  }
}

// True when every cell of the brick has the full 8th-order stencil in the
// backward pass: VOWR lowers the order only within 3 cells of the global
// gridphy boundary (vowr_order gives 4 for d >= 3, and always for R = 4).
// Block-uniform, so interior bricks run the constant-coefficient code.
template <int TX, int TY, int CY, int CZ>
__device__ __forceinline__ bool brick_full_order(const PhyDesc &D, int t, int gnx, int gny, int gnz, int R) {
  if (R >= 4) return true;
  const int tiles_x = (D.nx + TX - 1) / TX;
  const int tiles_y = (D.ny + TY * CY - 1) / (TY * CY);
  const int per = tiles_x * tiles_y;
  const int zc = t / per;
  const int rem = t - zc * per;
  const int ty = rem / tiles_x;
  const int tx = rem - ty * tiles_x;
  const int x0 = D.gx + tx * TX, x1 = D.gx + min(tx * TX + TX, D.nx) - 1;
  const int y0 = D.gy + ty * TY * CY, y1 = D.gy + min((ty + 1) * TY * CY, D.ny) - 1;
  const int z0 = D.gz + zc * CZ, z1 = D.gz + min(zc * CZ + CZ, D.nz) - 1;
  return x0 >= 3 && x1 <= gnx - 4 && y0 >= 3 && y1 <= gny - 4 && z0 >= 3 && z1 <= gnz - 4;
}

// Extra attributes of the brick kernels (occupancy experiments), empty by default.
#ifndef DDX_BRICK_ATTR
#define DDX_BRICK_ATTR
#endif

// XL: x arm from neighbouring lanes (tx = 64); every lane of a row then
// stays to the end (the shifts need the whole wavefront), and only lanes
// inside gridphy store.
template <int MODE, int TX, int TY, int CY, int CZ, class S, bool XL = false>
__global__ __launch_bounds__(TX *TY) DDX_BRICK_ATTR void k_vel_phy_zb(const PhyDesc *__restrict__ descs,
                                                       const int *__restrict__ starts, int n, float dx, float dy,
                                                       float dz, float sign, int gnx, int gny, int gnz, int R) {
  static_assert(!XL || TX == 64, "lane-exchange bricks need one wavefront per x row");
  const int b = brick_block(blockIdx.x, gridDim.x);
  const PhyDesc &D = descs[find_desc(starts, n, b)];
  gnx = D.gnx; gny = D.gny; gnz = D.gnz; R = D.R;  // per descriptor: fused batches mix grids
  int i, j0, k0;
  const bool live = zbrick<TX, TY, CY, CZ>(D, b - D.blk0, i, j0, k0);
  if (!XL && !live) return;
  // backward mode and reverse mode differ only in the coefficient rows
  if (MODE == RTM_BACKWARD && brick_full_order<TX, TY, CY, CZ>(D, b - D.blk0, gnx, gny, gnz, R))
    vel_zb_body<RTM_REVERSE, TX, TY, CY, CZ, S, XL>(D, i, j0, k0, dx, dy, dz, sign, gnx, gny, gnz, R, live);
  else
    vel_zb_body<MODE, TX, TY, CY, CZ, S, XL>(D, i, j0, k0, dx, dy, dz, sign, gnx, gny, gnz, R, live);
}

template <int MODE, int TX, int TY, int CY, int CZ, class S, bool XL = false>
__global__ __launch_bounds__(TX *TY) DDX_BRICK_ATTR void k_stress_phy_zb(const PhyDesc *__restrict__ descs,
                                                          const int *__restrict__ starts, int n, float dx,
                                                          float dy, float dz, float dt, int gnx, int gny, int gnz,
                                                          int R) {
  static_assert(!XL || TX == 64, "lane-exchange bricks need one wavefront per x row");
  const int b = brick_block(blockIdx.x, gridDim.x);
  const PhyDesc &D = descs[find_desc(starts, n, b)];
  gnx = D.gnx; gny = D.gny; gnz = D.gnz; R = D.R;  // per descriptor: fused batches mix grids
  int i, j0, k0;
  const bool live = zbrick<TX, TY, CY, CZ>(D, b - D.blk0, i, j0, k0);
  if (!XL && !live) return;
  if (MODE == RTM_BACKWARD && brick_full_order<TX, TY, CY, CZ>(D, b - D.blk0, gnx, gny, gnz, R))
    stress_zb_body<RTM_REVERSE, TX, TY, CY, CZ, S, XL>(D, i, j0, k0, dx, dy, dz, dt, gnx, gny, gnz, R, live);
  else
    stress_zb_body<MODE, TX, TY, CY, CZ, S, XL>(D, i, j0, k0, dx, dy, dz, dt, gnx, gny, gnz, R, live);
}

// ---------------------------------------------------------------------------
// Stress update with the y arm from LDS (YL bricks: tx = 64, one x row per
// wavefront, TY rows per block, CZ planes per thread).  The block's waves
// put the centre rows of vx, vy, vz (from their z register columns) and the
// 8 rows around the block (j0-4 .. j0-1, j0+TY .. j0+TY+3, loaded by the
// waves in turn) for all CZ planes into LDS, then one barrier, and every y
// tap of every plane comes from LDS.  The x arm comes from neighbouring
// lanes as in the XL bricks.  Same operands and order of operations as
// stress_zb_body (bit-identical).
// ---------------------------------------------------------------------------
template <int NF, int TY, int CZ>
struct YlTile {
  float v[NF][CZ][TY + 8][64];  // field, plane, row (j0-4 ..), x
};

template <int MODE, int TY, int CZ, class S>
__device__ __forceinline__ void stress_yl_body(YlTile<3, TY, CZ> &ly, const PhyDesc &D, int i, int j0, int k0,
                                               float dx, float dy, float dz, float dt, int gnx, int gny, int gnz,
                                               int R, bool live) {
  typedef typename GPtr<S>::T gT;
  const long sy = D.sy, sz = D.sz;
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
  const XLane xl = xlane(i, D.nx);
  const int ic = xl.ic;
  const int j = j0;        // zbrick's j0 is the thread's own row ...
  const int jt = j0 - w;   // ... and the tile starts w rows above it
  const int jc = min(j, D.ny + 3);  // rows past gridphy + 3 read a margin row and store nothing
  const gT *__restrict__ vx = (const gT *)D.f[VX];
  const gT *__restrict__ vy = (const gT *)D.f[VY];
  const gT *__restrict__ vz = (const gT *)D.f[VZ];
  const long base = (long)(D.ox + ic) + sy * (D.oy + jc) + sz * (D.oz + k0);
  // every load of the prologue is issued before the first LDS write, so the
  // tile costs one memory round trip: the wave's halo rows (h = w, w + TY,
  // ... < 8) into registers, then its z columns
  constexpr int HR = (8 + TY - 1) / TY;
  float hv[HR][CZ][3];
#pragma unroll
  for (int q = 0; q < HR; q++) {
    const int h = w + q * TY;
    const int rr = h < 4 ? h : TY + h;              // tile row
    const int jj = min(jt - 4 + rr, D.ny + 3);      // gridphy row, within the margin
    const long hb = (long)(D.ox + ic) + sy * (D.oy + jj) + sz * (D.oz + k0);
#pragma unroll
    for (int kk = 0; kk < CZ; kk++) {
      const bool use = h < 8 && k0 + kk < D.nz;
      hv[q][kk][0] = use ? ld(vx, hb + (long)kk * sz) : 0.f;
      hv[q][kk][1] = use ? ld(vy, hb + (long)kk * sz) : 0.f;
      hv[q][kk][2] = use ? ld(vz, hb + (long)kk * sz) : 0.f;
    }
  }
  float zx[CZ + 7], zy[CZ + 7], zz[CZ + 7];
  load_zcol<CZ>(zx, vx, base, sz, k0, -3, D.nz);
  load_zcol<CZ>(zy, vy, base, sz, k0, -3, D.nz);
  load_zcol<CZ>(zz, vz, base, sz, k0, -4, D.nz);
#pragma unroll
  for (int q = 0; q < HR; q++) {
    const int h = w + q * TY;
    if (h >= 8) break;
    const int rr = h < 4 ? h : TY + h;
#pragma unroll
    for (int kk = 0; kk < CZ; kk++) {
      ly.v[0][kk][rr][lane] = hv[q][kk][0];
      ly.v[1][kk][rr][lane] = hv[q][kk][1];
      ly.v[2][kk][rr][lane] = hv[q][kk][2];
    }
  }
#pragma unroll
  for (int kk = 0; kk < CZ; kk++) {
    ly.v[0][kk][w + 4][lane] = zx[kk + 3];
    ly.v[1][kk][w + 4][lane] = zy[kk + 3];
    ly.v[2][kk][w + 4][lane] = zz[kk + 4];
  }
  __syncthreads();
  if (j >= D.ny) return;  // whole wavefronts: nothing follows that needs them
  gT *__restrict__ fsxx = (gT *)D.f[SXX];
  gT *__restrict__ fsyy = (gT *)D.f[SYY];
  gT *__restrict__ fszz = (gT *)D.f[SZZ];
  gT *__restrict__ fsxy = (gT *)D.f[SXY];
  gT *__restrict__ fsxz = (gT *)D.f[SXZ];
  gT *__restrict__ fsyz = (gT *)D.f[SYZ];
  gT *__restrict__ fsau = (gT *)D.f[SAU];
  const gfloat *__restrict__ flam = G(D.lam);
  const gfloat *__restrict__ fmu = G(D.mu);
  const float a = dt / dx, bb = dt / dy, c = dt / dz;
  const Coef4 c4 = coef_row(4);
  const Coef4 cx = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gx + i, gnx, R)) : c4;
  const Coef4 cy = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gy + j, gny, R)) : c4;
  const long hbase = base + (xl.hx - ic);
  const int r0 = w + 4;  // the wave's row in the tile
#pragma unroll
  for (int kk = 0; kk < CZ; kk++) {
    const int k = k0 + kk;
    if (k >= D.nz) break;
    const Coef4 cz = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gz + k, gnz, R)) : c4;
    const long id = base + (long)kk * sz;
    const float oxx = ld_stream(fsxx, id), oyy = ld_stream(fsyy, id), ozz = ld_stream(fszz, id);
    const float oxy = ld_stream(fsxy, id), oxz = ld_stream(fsxz, id), oyz = ld_stream(fsyz, id);
    const float oau = rtm_mode(MODE) ? ld_stream(fsau, id) : 0.f;
    const float lm = ld_stream(flam, id), mu = ld_stream(fmu, id);
    const float l2m = lm + 2.0f * mu;
    const long hid = hbase + (long)kk * sz;
    const float gx = ld_halo(vx, xl, hid), gy = ld_halo(vy, xl, hid), gz = ld_halo(vz, xl, hid);
    const float vx_x = dmx(xrow(zx[kk + 3], gx), cx);
    const float vy_x = dpx(xrow(zy[kk + 3], gy), cx);
    const float vz_x = dpx(xrow(zz[kk + 4], gz), cx);
    const float(&tx)[TY + 8][64] = ly.v[0][kk];
    const float(&ty)[TY + 8][64] = ly.v[1][kk];
    const float(&tz)[TY + 8][64] = ly.v[2][kk];
    // forward-staggered in y: f(j+r+1) - f(j-r); backward: f(j+r) - f(j-r-1)
    float vx_y = 0.f;
    vx_y += cy.c0 * (tx[r0 + 1][lane] - tx[r0][lane]);
    vx_y += cy.c1 * (tx[r0 + 2][lane] - tx[r0 - 1][lane]);
    vx_y += cy.c2 * (tx[r0 + 3][lane] - tx[r0 - 2][lane]);
    vx_y += cy.c3 * (tx[r0 + 4][lane] - tx[r0 - 3][lane]);
    float vy_y = 0.f;
    vy_y += cy.c0 * (ty[r0][lane] - ty[r0 - 1][lane]);
    vy_y += cy.c1 * (ty[r0 + 1][lane] - ty[r0 - 2][lane]);
    vy_y += cy.c2 * (ty[r0 + 2][lane] - ty[r0 - 3][lane]);
    vy_y += cy.c3 * (ty[r0 + 3][lane] - ty[r0 - 4][lane]);
    float vz_y = 0.f;
    vz_y += cy.c0 * (tz[r0 + 1][lane] - tz[r0][lane]);
    vz_y += cy.c1 * (tz[r0 + 2][lane] - tz[r0 - 1][lane]);
    vz_y += cy.c2 * (tz[r0 + 3][lane] - tz[r0 - 2][lane]);
    vz_y += cy.c3 * (tz[r0 + 4][lane] - tz[r0 - 3][lane]);
    float vx_z = 0.f;
    vx_z += cz.c0 * (zx[kk + 4] - zx[kk + 3]);
    vx_z += cz.c1 * (zx[kk + 5] - zx[kk + 2]);
    vx_z += cz.c2 * (zx[kk + 6] - zx[kk + 1]);
    vx_z += cz.c3 * (zx[kk + 7] - zx[kk + 0]);
    float vy_z = 0.f;
    vy_z += cz.c0 * (zy[kk + 4] - zy[kk + 3]);
    vy_z += cz.c1 * (zy[kk + 5] - zy[kk + 2]);
    vy_z += cz.c2 * (zy[kk + 6] - zy[kk + 1]);
    vy_z += cz.c3 * (zy[kk + 7] - zy[kk + 0]);
    float vz_z = 0.f;
    vz_z += cz.c0 * (zz[kk + 4] - zz[kk + 3]);
    vz_z += cz.c1 * (zz[kk + 5] - zz[kk + 2]);
    vz_z += cz.c2 * (zz[kk + 6] - zz[kk + 1]);
    vz_z += cz.c3 * (zz[kk + 7] - zz[kk + 0]);
    if (!live) continue;
    const float nxx = oxx + (a * l2m * vx_x + bb * lm * vy_y + c * lm * vz_z);
    const float nyy = oyy + (a * lm * vx_x + bb * l2m * vy_y + c * lm * vz_z);
    const float nzz = ozz + (a * lm * vx_x + bb * lm * vy_y + c * l2m * vz_z);
    const float nxy = oxy + (a * mu * vy_x + bb * mu * vx_y);
    const float nxz = oxz + (a * mu * vz_x + c * mu * vx_z);
    const float nyz = oyz + (bb * mu * vz_y + c * mu * vy_z);
    const float nau = rtm_mode(MODE) ? oau + (a * l2m * vx_x + bb * l2m * vy_y + c * l2m * vz_z) : 0.f;
    st_stream(fsxx, id, nxx);
    st_stream(fsyy, id, nyy);
    st_stream(fszz, id, nzz);
    st_stream(fsxy, id, nxy);
    st_stream(fsxz, id, nxz);
    st_stream(fsyz, id, nyz);
    if constexpr (rtm_mode(MODE)) st_stream(fsau, id, nau);
    if (D.mir) {
      const MirCell m = mir_cell(D.mir, D.ox + i, D.oy + j, D.oz + k);
      if (m.any) {
        mir_put<gT>(m, SXX, nxx); mir_put<gT>(m, SYY, nyy); mir_put<gT>(m, SZZ, nzz);
        mir_put<gT>(m, SXY, nxy); mir_put<gT>(m, SXZ, nxz); mir_put<gT>(m, SYZ, nyz);
        if constexpr (rtm_mode(MODE)) mir_put<gT>(m, SAU, nau);
      }
    }
  }
}

template <int MODE, int TY, int CZ, class S>
__global__ __launch_bounds__(64 * TY) DDX_BRICK_ATTR void k_stress_phy_yl(const PhyDesc *__restrict__ descs,
                                                           const int *__restrict__ starts, int n, float dx,
                                                           float dy, float dz, float dt, int gnx, int gny,
                                                           int gnz, int R) {
  __shared__ YlTile<3, TY, CZ> ly;
  const int b = brick_block(blockIdx.x, gridDim.x);
  const PhyDesc &D = descs[find_desc(starts, n, b)];
  gnx = D.gnx; gny = D.gny; gnz = D.gnz; R = D.R;  // per descriptor: fused batches mix grids
  int i, j0, k0;
  const bool live = zbrick<64, TY, 1, CZ>(D, b - D.blk0, i, j0, k0);
  if (MODE == RTM_BACKWARD && brick_full_order<64, TY, 1, CZ>(D, b - D.blk0, gnx, gny, gnz, R))
    stress_yl_body<RTM_REVERSE, TY, CZ, S>(ly, D, i, j0, k0, dx, dy, dz, dt, gnx, gny, gnz, R, live);
  else
    stress_yl_body<MODE, TY, CZ, S>(ly, D, i, j0, k0, dx, dy, dz, dt, gnx, gny, gnz, R, live);
}

// Velocity update with the y arm from LDS: the tile holds syy, sxy, syz and,
// in the backward and reverse modes, sau (the velocity update's y-arm
// fields).  Same operands and order of operations as vel_zb_body.
template <int MODE, int TY, int CZ, class S>
__device__ __forceinline__ void vel_yl_body(YlTile<(MODE == RTM_BACKWARD || MODE == RTM_REVERSE) ? 4 : 3, TY, CZ> &ly,
                                            const PhyDesc &D, int i, int j0, int k0, float dx, float dy, float dz,
                                            float sign, int gnx, int gny, int gnz, int R, bool live) {
  typedef typename GPtr<S>::T gT;
  constexpr bool P = (MODE == RTM_BACKWARD || MODE == RTM_REVERSE);
  const long sy = D.sy, sz = D.sz;
  const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
  const XLane xl = xlane(i, D.nx);
  const int ic = xl.ic;
  const int j = j0;        // zbrick's j0 is the thread's own row ...
  const int jt = j0 - w;   // ... and the tile starts w rows above it
  const int jc = min(j, D.ny + 3);
  const gT *__restrict__ sxx = (const gT *)D.f[SXX];
  const gT *__restrict__ syy = (const gT *)D.f[SYY];
  const gT *__restrict__ szz = (const gT *)D.f[SZZ];
  const gT *__restrict__ sxy = (const gT *)D.f[SXY];
  const gT *__restrict__ sxz = (const gT *)D.f[SXZ];
  const gT *__restrict__ syz = (const gT *)D.f[SYZ];
  const gT *__restrict__ sau = (const gT *)D.f[SAU];
  const long base = (long)(D.ox + ic) + sy * (D.oy + jc) + sz * (D.oz + k0);
  // every load of the prologue is issued before the first LDS write (one
  // memory round trip): halo rows into registers, centre rows, z columns
  constexpr int NF = P ? 4 : 3;
  constexpr int HR = (8 + TY - 1) / TY;
  float hv[HR][CZ][NF];
#pragma unroll
  for (int q = 0; q < HR; q++) {
    const int h = w + q * TY;
    const int rr = h < 4 ? h : TY + h;
    const int jj = min(jt - 4 + rr, D.ny + 3);
    const long hb = (long)(D.ox + ic) + sy * (D.oy + jj) + sz * (D.oz + k0);
#pragma unroll
    for (int kk = 0; kk < CZ; kk++) {
      const bool use = h < 8 && k0 + kk < D.nz;
      hv[q][kk][0] = use ? ld(syy, hb + (long)kk * sz) : 0.f;
      hv[q][kk][1] = use ? ld(sxy, hb + (long)kk * sz) : 0.f;
      hv[q][kk][2] = use ? ld(syz, hb + (long)kk * sz) : 0.f;
      if constexpr (P) hv[q][kk][NF - 1] = use ? ld(sau, hb + (long)kk * sz) : 0.f;
    }
  }
  float cyy[CZ], cxy[CZ];  // planes k0 .. k0+CZ-1 lie within gridphy + 3
#pragma unroll
  for (int kk = 0; kk < CZ; kk++) {
    cyy[kk] = ld(syy, base + (long)kk * sz);
    cxy[kk] = ld(sxy, base + (long)kk * sz);
  }
  float zzz[CZ + 7], zxz[CZ + 7], zyz[CZ + 7], zsa[P ? CZ + 7 : 1];
  load_zcol<CZ>(zzz, szz, base, sz, k0, -3, D.nz);
  load_zcol<CZ>(zxz, sxz, base, sz, k0, -4, D.nz);
  load_zcol<CZ>(zyz, syz, base, sz, k0, -4, D.nz);
  if constexpr (P) load_zcol<CZ>(zsa, sau, base, sz, k0, -3, D.nz);
#pragma unroll
  for (int q = 0; q < HR; q++) {
    const int h = w + q * TY;
    if (h >= 8) break;
    const int rr = h < 4 ? h : TY + h;
#pragma unroll
    for (int kk = 0; kk < CZ; kk++) {
      ly.v[0][kk][rr][lane] = hv[q][kk][0];
      ly.v[1][kk][rr][lane] = hv[q][kk][1];
      ly.v[2][kk][rr][lane] = hv[q][kk][2];
      if constexpr (P) ly.v[P ? 3 : 0][kk][rr][lane] = hv[q][kk][NF - 1];
    }
  }
#pragma unroll
  for (int kk = 0; kk < CZ; kk++) {
    ly.v[0][kk][w + 4][lane] = cyy[kk];
    ly.v[1][kk][w + 4][lane] = cxy[kk];
    ly.v[2][kk][w + 4][lane] = zyz[kk + 4];
    if constexpr (P) ly.v[P ? 3 : 0][kk][w + 4][lane] = zsa[kk + 3];
  }
  __syncthreads();
  if (j >= D.ny) return;  // whole wavefronts
  gT *__restrict__ fvx = (gT *)D.f[VX];
  gT *__restrict__ fvy = (gT *)D.f[VY];
  gT *__restrict__ fvz = (gT *)D.f[VZ];
  gT *__restrict__ fvxp = (gT *)D.f[VXP];
  gT *__restrict__ fvyp = (gT *)D.f[VYP];
  gT *__restrict__ fvzp = (gT *)D.f[VZP];
  const gfloat *__restrict__ fdtr = G(D.dtr);
  const Coef4 c4 = coef_row(4);
  const Coef4 cx = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gx + i, gnx, R)) : c4;
  const Coef4 cy = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gy + j, gny, R)) : c4;
  const long hbase = base + (xl.hx - ic);
  const int r0 = w + 4;
#pragma unroll
  for (int kk = 0; kk < CZ; kk++) {
    const int k = k0 + kk;
    if (k >= D.nz) break;
    const Coef4 cz = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gz + k, gnz, R)) : c4;
    const long id = base + (long)kk * sz;
    const float ovx = ld_stream(fvx, id), ovy = ld_stream(fvy, id), ovz = ld_stream(fvz, id);
    const float ovxp = P ? ld_stream(fvxp, id) : 0.f, ovyp = P ? ld_stream(fvyp, id) : 0.f;
    const float ovzp = P ? ld_stream(fvzp, id) : 0.f;
    const float dtr = sign * ld_stream(fdtr, id);
    const float a = dtr / dx, bb = dtr / dy, c = dtr / dz;
    const long hid = hbase + (long)kk * sz;
    const float gxx = ld_halo(sxx, xl, hid), gxy = ld_halo(sxy, xl, hid), gxz = ld_halo(sxz, xl, hid);
    const float gsa = P ? ld_halo(sau, xl, hid) : 0.f;
    const float(&tyy)[TY + 8][64] = ly.v[0][kk];
    const float(&txy)[TY + 8][64] = ly.v[1][kk];
    const float(&tyz)[TY + 8][64] = ly.v[2][kk];
    const float sxx_x = dpx(xrow(ld(sxx, id), gxx), cx);
    float syy_y = 0.f;  // forward-staggered in y
    syy_y += cy.c0 * (tyy[r0 + 1][lane] - tyy[r0][lane]);
    syy_y += cy.c1 * (tyy[r0 + 2][lane] - tyy[r0 - 1][lane]);
    syy_y += cy.c2 * (tyy[r0 + 3][lane] - tyy[r0 - 2][lane]);
    syy_y += cy.c3 * (tyy[r0 + 4][lane] - tyy[r0 - 3][lane]);
    float szz_z = 0.f;
    szz_z += cz.c0 * (zzz[kk + 4] - zzz[kk + 3]);
    szz_z += cz.c1 * (zzz[kk + 5] - zzz[kk + 2]);
    szz_z += cz.c2 * (zzz[kk + 6] - zzz[kk + 1]);
    szz_z += cz.c3 * (zzz[kk + 7] - zzz[kk + 0]);
    const float syx_x = dmx(xrow(txy[r0][lane], gxy), cx);
    float sxy_y = 0.f;  // backward-staggered in y
    sxy_y += cy.c0 * (txy[r0][lane] - txy[r0 - 1][lane]);
    sxy_y += cy.c1 * (txy[r0 + 1][lane] - txy[r0 - 2][lane]);
    sxy_y += cy.c2 * (txy[r0 + 2][lane] - txy[r0 - 3][lane]);
    sxy_y += cy.c3 * (txy[r0 + 3][lane] - txy[r0 - 4][lane]);
    float sxz_z = 0.f;
    sxz_z += cz.c0 * (zxz[kk + 4] - zxz[kk + 3]);
    sxz_z += cz.c1 * (zxz[kk + 5] - zxz[kk + 2]);
    sxz_z += cz.c2 * (zxz[kk + 6] - zxz[kk + 1]);
    sxz_z += cz.c3 * (zxz[kk + 7] - zxz[kk + 0]);
    const float szx_x = dmx(xrow(zxz[kk + 4], gxz), cx);
    float szy_y = 0.f;
    szy_y += cy.c0 * (tyz[r0][lane] - tyz[r0 - 1][lane]);
    szy_y += cy.c1 * (tyz[r0 + 1][lane] - tyz[r0 - 2][lane]);
    szy_y += cy.c2 * (tyz[r0 + 2][lane] - tyz[r0 - 3][lane]);
    szy_y += cy.c3 * (tyz[r0 + 3][lane] - tyz[r0 - 4][lane]);
    float syz_z = 0.f;
    syz_z += cz.c0 * (zyz[kk + 4] - zyz[kk + 3]);
    syz_z += cz.c1 * (zyz[kk + 5] - zyz[kk + 2]);
    syz_z += cz.c2 * (zyz[kk + 6] - zyz[kk + 1]);
    syz_z += cz.c3 * (zyz[kk + 7] - zyz[kk + 0]);
    float sx = 0.f, sy_ = 0.f, sz_ = 0.f;
    if constexpr (P) {
      const float(&tsa)[TY + 8][64] = ly.v[P ? 3 : 0][kk];
      sx = dpx(xrow(zsa[kk + 3], gsa), cx);
      sy_ += cy.c0 * (tsa[r0 + 1][lane] - tsa[r0][lane]);
      sy_ += cy.c1 * (tsa[r0 + 2][lane] - tsa[r0 - 1][lane]);
      sy_ += cy.c2 * (tsa[r0 + 3][lane] - tsa[r0 - 2][lane]);
      sy_ += cy.c3 * (tsa[r0 + 4][lane] - tsa[r0 - 3][lane]);
      sz_ += cz.c0 * (zsa[kk + 4] - zsa[kk + 3]);
      sz_ += cz.c1 * (zsa[kk + 5] - zsa[kk + 2]);
      sz_ += cz.c2 * (zsa[kk + 6] - zsa[kk + 1]);
      sz_ += cz.c3 * (zsa[kk + 7] - zsa[kk + 0]);
    }
    if (!live) continue;
    const float nvx = ovx + (a * sxx_x + bb * sxy_y + c * sxz_z);
    const float nvy = ovy + (a * syx_x + bb * syy_y + c * syz_z);
    const float nvz = ovz + (a * szx_x + bb * szy_y + c * szz_z);
    st_stream(fvx, id, nvx);
    st_stream(fvy, id, nvy);
    st_stream(fvz, id, nvz);
    if (D.mir) {
      const MirCell m = mir_cell(D.mir, D.ox + i, D.oy + j, D.oz + k);
      if (m.any) { mir_put<gT>(m, VX, nvx); mir_put<gT>(m, VY, nvy); mir_put<gT>(m, VZ, nvz); }
    }
    if constexpr (P) {
      st_stream(fvxp, id, ovxp + a * sx);
      st_stream(fvyp, id, ovyp + bb * sy_);
      st_stream(fvzp, id, ovzp + c * sz_);
    }
  }
}

template <int MODE, int TY, int CZ, class S>
__global__ __launch_bounds__(64 * TY) DDX_BRICK_ATTR void k_vel_phy_yl(const PhyDesc *__restrict__ descs,
                                                        const int *__restrict__ starts, int n, float dx, float dy,
                                                        float dz, float sign, int gnx, int gny, int gnz, int R) {
  __shared__ YlTile<(MODE == RTM_BACKWARD || MODE == RTM_REVERSE) ? 4 : 3, TY, CZ> ly;
  const int b = brick_block(blockIdx.x, gridDim.x);
  const PhyDesc &D = descs[find_desc(starts, n, b)];
  gnx = D.gnx; gny = D.gny; gnz = D.gnz; R = D.R;  // per descriptor: fused batches mix grids
  int i, j0, k0;
  const bool live = zbrick<64, TY, 1, CZ>(D, b - D.blk0, i, j0, k0);
  // backward mode and reverse mode differ only in the coefficient rows
  if constexpr (MODE == RTM_BACKWARD) {
    if (brick_full_order<64, TY, 1, CZ>(D, b - D.blk0, gnx, gny, gnz, R))
      vel_yl_body<RTM_REVERSE, TY, CZ, S>(ly, D, i, j0, k0, dx, dy, dz, sign, gnx, gny, gnz, R, live);
    else
      vel_yl_body<MODE, TY, CZ, S>(ly, D, i, j0, k0, dx, dy, dz, sign, gnx, gny, gnz, R, live);
  } else {
    vel_yl_body<MODE, TY, CZ, S>(ly, D, i, j0, k0, dx, dy, dz, sign, gnx, gny, gnz, R, live);
  }
}

// ---------------------------------------------------------------------------
// Paired z-blocked bricks: every thread owns two neighbouring x cells
// (i, i + 1) with i even.  Rows start on 32-element boundaries (the
// storage layout), so every field access of the pair is one aligned
// 2-element load or store: 8 bytes per lane for FP32, 4 for the 16-bit
// formats.  That halves the memory instructions per cell, and keeps the
// bytes each instruction moves at the FP32 level for 16-bit storage.  The x
// stencils of both cells come from five pairs (cells i-4 .. i+5).  Operands
// and order of operations per cell are those of vel_zb_body /
// stress_zb_body (bit-identical results).  When gridphy has an odd width the
// last pair's second cell lies in the margin: it is read, never written.
// ---------------------------------------------------------------------------
typedef float f2v __attribute__((ext_vector_type(2)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) f2v gf2v;
typedef __attribute__((address_space(1))) h2v gh2v;
typedef __attribute__((address_space(1))) unsigned gu32;

struct F2 {
  float a, b;
};
__device__ __forceinline__ F2 ld2(const gfloat *p, long i) { const f2v v = *(const gf2v *)(p + i); return {v.x, v.y}; }
__device__ __forceinline__ F2 ld2(const ghalf *p, long i) {
  const h2v v = *(const gh2v *)(p + i);
  return {(float)v.x, (float)v.y};
}
__device__ __forceinline__ F2 ld2(const gbf16 *p, long i) {
  const unsigned u = *(const gu32 *)(p + i);
  return {__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u)};
}
__device__ __forceinline__ F2 ld2_nt(const gfloat *p, long i) {
  const f2v v = __builtin_nontemporal_load((const gf2v *)(p + i));
  return {v.x, v.y};
}
__device__ __forceinline__ F2 ld2_nt(const ghalf *p, long i) {
  const h2v v = __builtin_nontemporal_load((const gh2v *)(p + i));
  return {(float)v.x, (float)v.y};
}
__device__ __forceinline__ F2 ld2_nt(const gbf16 *p, long i) {
  const unsigned u = __builtin_nontemporal_load((const gu32 *)(p + i));
  return {__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u)};
}
// store the pair (both) or its first element only
__device__ __forceinline__ void st2_nt(gfloat *p, long i, float a, float b, bool both) {
  if (both) {
    f2v v;
    v.x = a;
    v.y = b;
    __builtin_nontemporal_store(v, (gf2v *)(p + i));
  } else {
    __builtin_nontemporal_store(a, p + i);
  }
}
__device__ __forceinline__ void st2_nt(ghalf *p, long i, float a, float b, bool both) {
  if (both) {
    h2v v;
    v.x = (_Float16)a;
    v.y = (_Float16)b;
    __builtin_nontemporal_store(v, (gh2v *)(p + i));
  } else {
    __builtin_nontemporal_store((_Float16)a, p + i);
  }
}
__device__ __forceinline__ void st2_nt(gbf16 *p, long i, float a, float b, bool both) {
  if (both) {
    __builtin_nontemporal_store((unsigned)bf16_rne(a) | ((unsigned)bf16_rne(b) << 16), (gu32 *)(p + i));
  } else {
    __builtin_nontemporal_store(bf16_rne(a), p + i);
  }
}
template <typename GT>
__device__ __forceinline__ void rmw2(GT *p, long i, float va, float vb, bool both) {
  const F2 o = ld2_nt(p, i);
  st2_nt(p, i, o.a + va, o.b + vb, both);
}

// cells i-4 .. i+5 of a row (i even): v[k] = f(i - 4 + k)
template <typename GT>
__device__ __forceinline__ void ld_xwin(float (&v)[10], const GT *f, long id) {
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const F2 q = ld2(f, id - 4 + 2 * m);
    v[2 * m] = q.a;
    v[2 * m + 1] = q.b;
  }
}
// forward-staggered x derivative of cells i (a) and i+1 (b) from the window
__device__ __forceinline__ F2 dpx2(const float (&v)[10], const Coef4 &ca, const Coef4 &cb) {
  float a = 0.f, b = 0.f;
  a += ca.c0 * (v[5] - v[4]);
  a += ca.c1 * (v[6] - v[3]);
  a += ca.c2 * (v[7] - v[2]);
  a += ca.c3 * (v[8] - v[1]);
  b += cb.c0 * (v[6] - v[5]);
  b += cb.c1 * (v[7] - v[4]);
  b += cb.c2 * (v[8] - v[3]);
  b += cb.c3 * (v[9] - v[2]);
  return {a, b};
}
// backward-staggered x derivative
__device__ __forceinline__ F2 dmx2(const float (&v)[10], const Coef4 &ca, const Coef4 &cb) {
  float a = 0.f, b = 0.f;
  a += ca.c0 * (v[4] - v[3]);
  a += ca.c1 * (v[5] - v[2]);
  a += ca.c2 * (v[6] - v[1]);
  a += ca.c3 * (v[7] - v[0]);
  b += cb.c0 * (v[5] - v[4]);
  b += cb.c1 * (v[6] - v[3]);
  b += cb.c2 * (v[7] - v[2]);
  b += cb.c3 * (v[8] - v[1]);
  return {a, b};
}
// y (or any stride) derivatives of both cells from pair loads
template <typename GT>
__device__ __forceinline__ F2 dp4_2(const GT *f, long i, long s, const Coef4 &c) {
  float a = 0.f, b = 0.f;
  F2 p = ld2(f, i + s), q = ld2(f, i);
  a += c.c0 * (p.a - q.a); b += c.c0 * (p.b - q.b);
  p = ld2(f, i + 2 * s); q = ld2(f, i - s);
  a += c.c1 * (p.a - q.a); b += c.c1 * (p.b - q.b);
  p = ld2(f, i + 3 * s); q = ld2(f, i - 2 * s);
  a += c.c2 * (p.a - q.a); b += c.c2 * (p.b - q.b);
  p = ld2(f, i + 4 * s); q = ld2(f, i - 3 * s);
  a += c.c3 * (p.a - q.a); b += c.c3 * (p.b - q.b);
  return {a, b};
}
template <typename GT>
__device__ __forceinline__ F2 dm4_2(const GT *f, long i, long s, const Coef4 &c) {
  float a = 0.f, b = 0.f;
  F2 p = ld2(f, i), q = ld2(f, i - s);
  a += c.c0 * (p.a - q.a); b += c.c0 * (p.b - q.b);
  p = ld2(f, i + s); q = ld2(f, i - 2 * s);
  a += c.c1 * (p.a - q.a); b += c.c1 * (p.b - q.b);
  p = ld2(f, i + 2 * s); q = ld2(f, i - 3 * s);
  a += c.c2 * (p.a - q.a); b += c.c2 * (p.b - q.b);
  p = ld2(f, i + 3 * s); q = ld2(f, i - 4 * s);
  a += c.c3 * (p.a - q.a); b += c.c3 * (p.b - q.b);
  return {a, b};
}
// z column of the pair: za/zb[r] = f(plane k0 + lo + r), planes past nz + 3 read as 0
template <int CZ, typename GT>
__device__ __forceinline__ void load_zcol2(float (&za)[CZ + 7], float (&zb)[CZ + 7], const GT *__restrict__ f,
                                           long base, long sz, int k0, int lo, int nz) {
#pragma unroll
  for (int r = 0; r < CZ + 7; r++) {
    if (k0 + lo + r < nz + 4) {
      const F2 q = ld2(f, base + (long)(lo + r) * sz);
      za[r] = q.a;
      zb[r] = q.b;
    } else {
      za[r] = 0.f;
      zb[r] = 0.f;
    }
  }
}
// z derivative from a column: forward- and backward-staggered columns are
// loaded from planes k-3 / k-4 so both read zc[kk+4+r] - zc[kk+3-r]
template <int N>
__device__ __forceinline__ float dzq(const float (&zc)[N], int kk, const Coef4 &c) {
  float a = 0.f;
  a += c.c0 * (zc[kk + 4] - zc[kk + 3]);
  a += c.c1 * (zc[kk + 5] - zc[kk + 2]);
  a += c.c2 * (zc[kk + 6] - zc[kk + 1]);
  a += c.c3 * (zc[kk + 7] - zc[kk + 0]);
  return a;
}

// brick of pairs: TX threads cover 2*TX cells in x
template <int TX, int TY, int CY, int CZ>
__device__ __forceinline__ bool zbrick2(const PhyDesc &D, int t, int &i, int &j0, int &k0) {
  const int tiles_x = (D.nx + 2 * TX - 1) / (2 * TX);
  const int tiles_y = (D.ny + TY * CY - 1) / (TY * CY);
  const int per = tiles_x * tiles_y;
  const int zc = t / per;
  const int rem = t - zc * per;
  const int ty = rem / tiles_x;
  const int tx = rem - ty * tiles_x;
  i = tx * 2 * TX + 2 * (int)(threadIdx.x % TX);
  j0 = ty * TY * CY + (int)(threadIdx.x / TX);
  k0 = zc * CZ;
  return i < D.nx;
}
template <int TX, int TY, int CY, int CZ>
__device__ __forceinline__ bool brick2_full_order(const PhyDesc &D, int t, int gnx, int gny, int gnz, int R) {
  if (R >= 4) return true;
  const int tiles_x = (D.nx + 2 * TX - 1) / (2 * TX);
  const int tiles_y = (D.ny + TY * CY - 1) / (TY * CY);
  const int per = tiles_x * tiles_y;
  const int zc = t / per;
  const int rem = t - zc * per;
  const int ty = rem / tiles_x;
  const int tx = rem - ty * tiles_x;
  const int x0 = D.gx + tx * 2 * TX, x1 = D.gx + min(tx * 2 * TX + 2 * TX, D.nx) - 1;
  const int y0 = D.gy + ty * TY * CY, y1 = D.gy + min((ty + 1) * TY * CY, D.ny) - 1;
  const int z0 = D.gz + zc * CZ, z1 = D.gz + min(zc * CZ + CZ, D.nz) - 1;
  return x0 >= 3 && x1 <= gnx - 4 && y0 >= 3 && y1 <= gny - 4 && z0 >= 3 && z1 <= gnz - 4;
}

template <int MODE, int TX, int TY, int CY, int CZ, class S>
__device__ __forceinline__ void vel_zb2_body(const PhyDesc &D, int i, int j0, int k0, float dx, float dy, float dz,
                                             float sign, int gnx, int gny, int gnz, int R) {
  static_assert(MODE != FWI, "the FWI mode runs on the one-cell bricks");
  typedef typename GPtr<S>::T gT;
  const long sy = D.sy, sz = D.sz;
  const gT *__restrict__ sxx = (const gT *)D.f[SXX];
  const gT *__restrict__ syy = (const gT *)D.f[SYY];
  const gT *__restrict__ szz = (const gT *)D.f[SZZ];
  const gT *__restrict__ sxy = (const gT *)D.f[SXY];
  const gT *__restrict__ sxz = (const gT *)D.f[SXZ];
  const gT *__restrict__ syz = (const gT *)D.f[SYZ];
  const gT *__restrict__ sau = (const gT *)D.f[SAU];
  gT *__restrict__ fvx = (gT *)D.f[VX];
  gT *__restrict__ fvy = (gT *)D.f[VY];
  gT *__restrict__ fvz = (gT *)D.f[VZ];
  gT *__restrict__ fvxp = (gT *)D.f[VXP];
  gT *__restrict__ fvyp = (gT *)D.f[VYP];
  gT *__restrict__ fvzp = (gT *)D.f[VZP];
  const gfloat *__restrict__ fdtr = G(D.dtr);
  constexpr bool P = (MODE == RTM_BACKWARD || MODE == RTM_REVERSE);
  const bool both = i + 1 < D.nx;
  const Coef4 c4 = coef_row(4);
  const Coef4 cxa = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gx + i, gnx, R)) : c4;
  const Coef4 cxb = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gx + i + 1, gnx, R)) : c4;
#pragma unroll
  for (int q = 0; q < CY; q++) {
    const int j = j0 + q * TY;
    if (j >= D.ny) continue;
    const Coef4 cy = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gy + j, gny, R)) : c4;
    const long base = (long)(D.ox + i) + sy * (D.oy + j) + sz * (D.oz + k0);
    float zzza[CZ + 7], zzzb[CZ + 7], zxza[CZ + 7], zxzb[CZ + 7], zyza[CZ + 7], zyzb[CZ + 7];
    float zsaa[P ? CZ + 7 : 1], zsab[P ? CZ + 7 : 1];
    load_zcol2<CZ>(zzza, zzzb, szz, base, sz, k0, -3, D.nz);
    load_zcol2<CZ>(zxza, zxzb, sxz, base, sz, k0, -4, D.nz);
    load_zcol2<CZ>(zyza, zyzb, syz, base, sz, k0, -4, D.nz);
    if constexpr (P) load_zcol2<CZ>(zsaa, zsab, sau, base, sz, k0, -3, D.nz);
#pragma unroll
    for (int kk = 0; kk < CZ; kk++) {
      const int k = k0 + kk;
      if (k >= D.nz) break;
      const Coef4 cz = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gz + k, gnz, R)) : c4;
      const long id = base + (long)kk * sz;
      const F2 ovx = ld2_nt(fvx, id), ovy = ld2_nt(fvy, id), ovz = ld2_nt(fvz, id);
      F2 ovxp{0.f, 0.f}, ovyp{0.f, 0.f}, ovzp{0.f, 0.f};
      if constexpr (P) {
        ovxp = ld2_nt(fvxp, id);
        ovyp = ld2_nt(fvyp, id);
        ovzp = ld2_nt(fvzp, id);
      }
      const F2 dr = ld2_nt(fdtr, id);
      const float dtra = sign * dr.a, dtrb = sign * dr.b;
      const float aa = dtra / dx, ba = dtra / dy, ca = dtra / dz;
      const float ab = dtrb / dx, bb = dtrb / dy, cb = dtrb / dz;
      float w[10];
      ld_xwin(w, sxx, id);
      const F2 sxx_x = dpx2(w, cxa, cxb);
      const F2 syy_y = dp4_2(syy, id, sy, cy);
      const float szz_za = dzq(zzza, kk, cz), szz_zb = dzq(zzzb, kk, cz);
      ld_xwin(w, sxy, id);
      const F2 syx_x = dmx2(w, cxa, cxb);
      const F2 sxy_y = dm4_2(sxy, id, sy, cy);
      const float sxz_za = dzq(zxza, kk, cz), sxz_zb = dzq(zxzb, kk, cz);
      ld_xwin(w, sxz, id);
      const F2 szx_x = dmx2(w, cxa, cxb);
      const F2 szy_y = dm4_2(syz, id, sy, cy);
      const float syz_za = dzq(zyza, kk, cz), syz_zb = dzq(zyzb, kk, cz);
      F2 sx{0.f, 0.f}, sy_{0.f, 0.f};
      float sza = 0.f, szb = 0.f;
      if constexpr (P) {
        ld_xwin(w, sau, id);
        sx = dpx2(w, cxa, cxb);
        sy_ = dp4_2(sau, id, sy, cy);
        sza = dzq(zsaa, kk, cz);
        szb = dzq(zsab, kk, cz);
      }
      const float nxa = ovx.a + (aa * sxx_x.a + ba * sxy_y.a + ca * sxz_za);
      const float nxb = ovx.b + (ab * sxx_x.b + bb * sxy_y.b + cb * sxz_zb);
      const float nya = ovy.a + (aa * syx_x.a + ba * syy_y.a + ca * syz_za);
      const float nyb = ovy.b + (ab * syx_x.b + bb * syy_y.b + cb * syz_zb);
      const float nza = ovz.a + (aa * szx_x.a + ba * szy_y.a + ca * szz_za);
      const float nzb = ovz.b + (ab * szx_x.b + bb * szy_y.b + cb * szz_zb);
      st2_nt(fvx, id, nxa, nxb, both);
      st2_nt(fvy, id, nya, nyb, both);
      st2_nt(fvz, id, nza, nzb, both);
      if (D.mir) {
        const MirCell ma = mir_cell(D.mir, D.ox + i, D.oy + j, D.oz + k);
        const MirCell mb = mir_cell(D.mir, D.ox + i + 1, D.oy + j, D.oz + k);
        mir_pair<gT>(ma, mb, both, VX, nxa, nxb);
        mir_pair<gT>(ma, mb, both, VY, nya, nyb);
        mir_pair<gT>(ma, mb, both, VZ, nza, nzb);
      }
      if constexpr (P) {
        st2_nt(fvxp, id, ovxp.a + aa * sx.a, ovxp.b + ab * sx.b, both);
        st2_nt(fvyp, id, ovyp.a + ba * sy_.a, ovyp.b + bb * sy_.b, both);
        st2_nt(fvzp, id, ovzp.a + ca * sza, ovzp.b + cb * szb, both);
      }
    }
  }
}

template <int MODE, int TX, int TY, int CY, int CZ, class S>
__device__ __forceinline__ void stress_zb2_body(const PhyDesc &D, int i, int j0, int k0, float dx, float dy,
                                                float dz, float dt, int gnx, int gny, int gnz, int R) {
  static_assert(MODE != FWI, "the FWI mode runs on the one-cell bricks");
  typedef typename GPtr<S>::T gT;
  const long sy = D.sy, sz = D.sz;
  const gT *__restrict__ vx = (const gT *)D.f[VX];
  const gT *__restrict__ vy = (const gT *)D.f[VY];
  const gT *__restrict__ vz = (const gT *)D.f[VZ];
  gT *__restrict__ fsxx = (gT *)D.f[SXX];
  gT *__restrict__ fsyy = (gT *)D.f[SYY];
  gT *__restrict__ fszz = (gT *)D.f[SZZ];
  gT *__restrict__ fsxy = (gT *)D.f[SXY];
  gT *__restrict__ fsxz = (gT *)D.f[SXZ];
  gT *__restrict__ fsyz = (gT *)D.f[SYZ];
  gT *__restrict__ fsau = (gT *)D.f[SAU];
  const gfloat *__restrict__ flam = G(D.lam);
  const gfloat *__restrict__ fmu = G(D.mu);
  const float a = dt / dx, bb = dt / dy, c = dt / dz;
  const bool both = i + 1 < D.nx;
  const Coef4 c4 = coef_row(4);
  const Coef4 cxa = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gx + i, gnx, R)) : c4;
  const Coef4 cxb = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gx + i + 1, gnx, R)) : c4;
#pragma unroll
  for (int q = 0; q < CY; q++) {
    const int j = j0 + q * TY;
    if (j >= D.ny) continue;
    const Coef4 cy = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gy + j, gny, R)) : c4;
    const long base = (long)(D.ox + i) + sy * (D.oy + j) + sz * (D.oz + k0);
    float zxa[CZ + 7], zxb[CZ + 7], zya[CZ + 7], zyb[CZ + 7], zza[CZ + 7], zzb[CZ + 7];
    load_zcol2<CZ>(zxa, zxb, vx, base, sz, k0, -3, D.nz);  // vx, vy: forward-staggered in z
    load_zcol2<CZ>(zya, zyb, vy, base, sz, k0, -3, D.nz);
    load_zcol2<CZ>(zza, zzb, vz, base, sz, k0, -4, D.nz);  // vz: backward-staggered in z
#pragma unroll
    for (int kk = 0; kk < CZ; kk++) {
      const int k = k0 + kk;
      if (k >= D.nz) break;
      const Coef4 cz = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gz + k, gnz, R)) : c4;
      const long id = base + (long)kk * sz;
      const F2 oxx = ld2_nt(fsxx, id), oyy = ld2_nt(fsyy, id), ozz = ld2_nt(fszz, id);
      const F2 oxy = ld2_nt(fsxy, id), oxz = ld2_nt(fsxz, id), oyz = ld2_nt(fsyz, id);
      F2 oau{0.f, 0.f};
      if constexpr (rtm_mode(MODE)) oau = ld2_nt(fsau, id);
      const F2 lm = ld2_nt(flam, id), mu = ld2_nt(fmu, id);
      const float l2ma = lm.a + 2.0f * mu.a, l2mb = lm.b + 2.0f * mu.b;
      float w[10];
      ld_xwin(w, vx, id);
      const F2 vx_x = dmx2(w, cxa, cxb);
      const F2 vx_y = dp4_2(vx, id, sy, cy);
      const float vx_za = dzq(zxa, kk, cz), vx_zb = dzq(zxb, kk, cz);
      ld_xwin(w, vy, id);
      const F2 vy_x = dpx2(w, cxa, cxb);
      const F2 vy_y = dm4_2(vy, id, sy, cy);
      const float vy_za = dzq(zya, kk, cz), vy_zb = dzq(zyb, kk, cz);
      ld_xwin(w, vz, id);
      const F2 vz_x = dpx2(w, cxa, cxb);
      const F2 vz_y = dp4_2(vz, id, sy, cy);
      const float vz_za = dzq(zza, kk, cz), vz_zb = dzq(zzb, kk, cz);
      const float nxxa = oxx.a + (a * l2ma * vx_x.a + bb * lm.a * vy_y.a + c * lm.a * vz_za);
      const float nxxb = oxx.b + (a * l2mb * vx_x.b + bb * lm.b * vy_y.b + c * lm.b * vz_zb);
      const float nyya = oyy.a + (a * lm.a * vx_x.a + bb * l2ma * vy_y.a + c * lm.a * vz_za);
      const float nyyb = oyy.b + (a * lm.b * vx_x.b + bb * l2mb * vy_y.b + c * lm.b * vz_zb);
      const float nzza = ozz.a + (a * lm.a * vx_x.a + bb * lm.a * vy_y.a + c * l2ma * vz_za);
      const float nzzb = ozz.b + (a * lm.b * vx_x.b + bb * lm.b * vy_y.b + c * l2mb * vz_zb);
      const float nxya = oxy.a + (a * mu.a * vy_x.a + bb * mu.a * vx_y.a);
      const float nxyb = oxy.b + (a * mu.b * vy_x.b + bb * mu.b * vx_y.b);
      const float nxza = oxz.a + (a * mu.a * vz_x.a + c * mu.a * vx_za);
      const float nxzb = oxz.b + (a * mu.b * vz_x.b + c * mu.b * vx_zb);
      const float nyza = oyz.a + (bb * mu.a * vz_y.a + c * mu.a * vy_za);
      const float nyzb = oyz.b + (bb * mu.b * vz_y.b + c * mu.b * vy_zb);
      const float naua = rtm_mode(MODE) ? oau.a + (a * l2ma * vx_x.a + bb * l2ma * vy_y.a + c * l2ma * vz_za) : 0.f;
      const float naub = rtm_mode(MODE) ? oau.b + (a * l2mb * vx_x.b + bb * l2mb * vy_y.b + c * l2mb * vz_zb) : 0.f;
      st2_nt(fsxx, id, nxxa, nxxb, both);
      st2_nt(fsyy, id, nyya, nyyb, both);
      st2_nt(fszz, id, nzza, nzzb, both);
      st2_nt(fsxy, id, nxya, nxyb, both);
      st2_nt(fsxz, id, nxza, nxzb, both);
      st2_nt(fsyz, id, nyza, nyzb, both);
      if constexpr (rtm_mode(MODE)) st2_nt(fsau, id, naua, naub, both);
      if (D.mir) {
        const MirCell ma = mir_cell(D.mir, D.ox + i, D.oy + j, D.oz + k);
        const MirCell mb = mir_cell(D.mir, D.ox + i + 1, D.oy + j, D.oz + k);
        mir_pair<gT>(ma, mb, both, SXX, nxxa, nxxb);
        mir_pair<gT>(ma, mb, both, SYY, nyya, nyyb);
        mir_pair<gT>(ma, mb, both, SZZ, nzza, nzzb);
        mir_pair<gT>(ma, mb, both, SXY, nxya, nxyb);
        mir_pair<gT>(ma, mb, both, SXZ, nxza, nxzb);
        mir_pair<gT>(ma, mb, both, SYZ, nyza, nyzb);
        if constexpr (rtm_mode(MODE)) mir_pair<gT>(ma, mb, both, SAU, naua, naub);
      }
    }
  }
}

template <int MODE, int TX, int TY, int CY, int CZ, class S>
__global__ __launch_bounds__(TX *TY) DDX_BRICK_ATTR void k_vel_phy_zb2(const PhyDesc *__restrict__ descs,
                                                        const int *__restrict__ starts, int n, float dx, float dy,
                                                        float dz, float sign, int gnx, int gny, int gnz, int R) {
  const int b = brick_block(blockIdx.x, gridDim.x);
  const PhyDesc &D = descs[find_desc(starts, n, b)];
  gnx = D.gnx; gny = D.gny; gnz = D.gnz; R = D.R;  // per descriptor: fused batches mix grids
  int i, j0, k0;
  if (!zbrick2<TX, TY, CY, CZ>(D, b - D.blk0, i, j0, k0)) return;
  if (MODE == RTM_BACKWARD && brick2_full_order<TX, TY, CY, CZ>(D, b - D.blk0, gnx, gny, gnz, R))
    vel_zb2_body<RTM_REVERSE, TX, TY, CY, CZ, S>(D, i, j0, k0, dx, dy, dz, sign, gnx, gny, gnz, R);
  else
    vel_zb2_body<MODE, TX, TY, CY, CZ, S>(D, i, j0, k0, dx, dy, dz, sign, gnx, gny, gnz, R);
}

template <int MODE, int TX, int TY, int CY, int CZ, class S>
__global__ __launch_bounds__(TX *TY) DDX_BRICK_ATTR void k_stress_phy_zb2(const PhyDesc *__restrict__ descs,
                                                           const int *__restrict__ starts, int n, float dx,
                                                           float dy, float dz, float dt, int gnx, int gny, int gnz,
                                                           int R) {
  const int b = brick_block(blockIdx.x, gridDim.x);
  const PhyDesc &D = descs[find_desc(starts, n, b)];
  gnx = D.gnx; gny = D.gny; gnz = D.gnz; R = D.R;  // per descriptor: fused batches mix grids
  int i, j0, k0;
  if (!zbrick2<TX, TY, CY, CZ>(D, b - D.blk0, i, j0, k0)) return;
  if (MODE == RTM_BACKWARD && brick2_full_order<TX, TY, CY, CZ>(D, b - D.blk0, gnx, gny, gnz, R))
    stress_zb2_body<RTM_REVERSE, TX, TY, CY, CZ, S>(D, i, j0, k0, dx, dy, dz, dt, gnx, gny, gnz, R);
  else
    stress_zb2_body<MODE, TX, TY, CY, CZ, S>(D, i, j0, k0, dx, dy, dz, dt, gnx, gny, gnz, R);
}

// ---------------------------------------------------------------------------
// LDS-tiled 2.5D versions of the physical-region updates (selectable
// alternative to the z-blocked bricks).  A block of 64 x 8 threads owns a 64 x 8 column tile and marches a
// run of kz planes.  For every plane, the fields differentiated along x or y
// are staged in LDS with their 4-cell halos (double-buffered, one barrier
// per plane); the z arms stay in per-thread register queues of 8 planes.
// Loads of plane k+1 are issued before plane k is computed.  Per cell this
// costs about 14 global loads instead of about 50 for the one-thread-per-tap
// form, and the taps are served by LDS.  Operands and order of operations
// are those of k_vel_phy / k_stress_phy (bit-identical results).
// ---------------------------------------------------------------------------
constexpr int LX = 64, LY = 8, HXW = LX + 8, HYW = LY + 8;

struct LdsMap {
  int i, j;           // this thread's column (subdomain gridphy coordinates)
  int hrow;           // y-halo row this thread loads (gridphy y), LDS row hr
  int hr;
  int xr, xc, xcol;   // x-halo (wave 0 only): LDS row xr, LDS column xc, gridphy x xcol
  bool ld;            // this thread's column lies in the loadable region
  bool hld, xld;      // halo loads in range
  bool comp;          // computes an output cell
};

__device__ __forceinline__ LdsMap lds_map(const PhyDesc &D, int x0, int y0) {
  LdsMap m;
  const int tx = threadIdx.x % LX, ty = threadIdx.x / LX;
  m.i = x0 + tx;
  m.j = y0 + ty;
  m.ld = m.i < D.nx + 4 && m.j < D.ny + 4;
  m.comp = m.i < D.nx && m.j < D.ny;
  m.hr = ty < 4 ? ty : ty + LY;
  m.hrow = ty < 4 ? y0 - 4 + ty : y0 + LY + (ty - 4);
  m.hld = m.i < D.nx + 4 && m.hrow < D.ny + 4;
  const int hx = tx & 7;
  m.xr = tx >> 3;
  m.xc = hx < 4 ? hx : LX + hx;
  m.xcol = hx < 4 ? x0 - 4 + hx : x0 + LX + (hx - 4);
  m.xld = ty == 0 && m.xcol < D.nx + 4 && (y0 + m.xr) < D.ny + 4;
  return m;
}

template <int MODE>
__global__ __launch_bounds__(LX *LY) void k_vel_lds(const PhyDesc *__restrict__ descs,
                                                    const int *__restrict__ starts, int n, float dx, float dy,
                                                    float dz, float sign, int gnx, int gny, int gnz, int R,
                                                    int kz) {
  constexpr bool P = (MODE == RTM_BACKWARD || MODE == RTM_REVERSE);
  __shared__ float Lxx[2][LY][HXW];
  __shared__ float Lyy[2][HYW][LX];
  __shared__ float Lxy[2][HYW][HXW];
  __shared__ float Lxz[2][LY][HXW];
  __shared__ float Lyz[2][HYW][LX];
  __shared__ float Lsa[P ? 2 : 1][P ? HYW : 1][P ? HXW : 1];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const PhyDesc &D = descs[find_desc(starts, n, b)];
  gnx = D.gnx; gny = D.gny; gnz = D.gnz; R = D.R;  // per descriptor: fused batches mix grids
  const int t = b - D.blk0;
  const int tiles_x = (D.nx + LX - 1) / LX, tiles_y = (D.ny + LY - 1) / LY, per = tiles_x * tiles_y;
  const int zc = t / per, rem = t - zc * per, tyi = rem / tiles_x, txi = rem - tyi * tiles_x;
  const int x0 = txi * LX, y0 = tyi * LY, k0 = zc * kz, k1 = min(k0 + kz, D.nz);
  const LdsMap m = lds_map(D, x0, y0);
  const int tx = threadIdx.x % LX, ty = threadIdx.x / LX;
  const long sy = D.sy, sz = D.sz;
  const gfloat *sxx = G(D.f[SXX]), *syy = G(D.f[SYY]), *szz = G(D.f[SZZ]);
  const gfloat *sxy = G(D.f[SXY]), *sxz = G(D.f[SXZ]), *syz = G(D.f[SYZ]), *sau = G(D.f[SAU]);
  const long cb = (long)(D.ox + m.i) + sy * (D.oy + m.j) + sz * D.oz;        // column base (plane 0)
  const long hb = (long)(D.ox + m.i) + sy * (D.oy + m.hrow) + sz * D.oz;     // y-halo base
  const long xb = (long)(D.ox + m.xcol) + sy * (D.oy + y0 + m.xr) + sz * D.oz;  // x-halo base
  const Coef4 c4 = coef_row(4);
  const Coef4 cx = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gx + min(m.i, D.nx - 1), gnx, R)) : c4;
  const Coef4 cy = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gy + min(m.j, D.ny - 1), gny, R)) : c4;
  // z-arm queues: qzz, qsa planes k-3..k+4; qxz, qyz planes k-4..k+3
  float qzz[8], qxz[8], qyz[8], qsa[8];
#pragma unroll
  for (int r = 0; r < 8; r++) {
    qzz[r] = m.ld ? szz[cb + (long)(k0 - 3 + r) * sz] : 0.f;
    qxz[r] = m.ld ? sxz[cb + (long)(k0 - 4 + r) * sz] : 0.f;
    qyz[r] = m.ld ? syz[cb + (long)(k0 - 4 + r) * sz] : 0.f;
    if (P) qsa[r] = m.ld ? sau[cb + (long)(k0 - 3 + r) * sz] : 0.f;
  }
  // current-plane staging registers
  float c_xx = 0.f, c_yy = 0.f, c_xy = 0.f;          // interior
  float h_yy = 0.f, h_xy = 0.f, h_yz = 0.f, h_sa = 0.f;  // y halo
  float x_xx = 0.f, x_xy = 0.f, x_xz = 0.f, x_sa = 0.f;  // x halo
  auto load_plane = [&](int k, float &a_xx, float &a_yy, float &a_xy, float &b_yy, float &b_xy, float &b_yz,
                        float &b_sa, float &e_xx, float &e_xy, float &e_xz, float &e_sa) {
    const long kk = (long)k * sz;
    if (m.ld) { a_xx = sxx[cb + kk]; a_yy = syy[cb + kk]; a_xy = sxy[cb + kk]; }
    if (m.hld) {
      b_yy = syy[hb + kk]; b_xy = sxy[hb + kk]; b_yz = syz[hb + kk];
      if (P) b_sa = sau[hb + kk];
    }
    if (m.xld) {
      e_xx = sxx[xb + kk]; e_xy = sxy[xb + kk]; e_xz = sxz[xb + kk];
      if (P) e_sa = sau[xb + kk];
    }
  };
  load_plane(k0, c_xx, c_yy, c_xy, h_yy, h_xy, h_yz, h_sa, x_xx, x_xy, x_xz, x_sa);
  gfloat *fvx = G(D.f[VX]), *fvy = G(D.f[VY]), *fvz = G(D.f[VZ]);
  gfloat *fvxp = G(D.f[VXP]), *fvyp = G(D.f[VYP]), *fvzp = G(D.f[VZP]);
  const gfloat *fdtr = G(D.dtr);
  for (int k = k0; k < k1; k++) {
    const int bf = (k - k0) & 1;
    Lxx[bf][ty][tx + 4] = c_xx;
    Lyy[bf][ty + 4][tx] = c_yy;
    Lxy[bf][ty + 4][tx + 4] = c_xy;
    Lxz[bf][ty][tx + 4] = qxz[4];
    Lyz[bf][ty + 4][tx] = qyz[4];
    if (P) Lsa[bf][ty + 4][tx + 4] = qsa[3];
    Lyy[bf][m.hr][tx] = h_yy;
    Lxy[bf][m.hr][tx + 4] = h_xy;
    Lyz[bf][m.hr][tx] = h_yz;
    if (P) Lsa[bf][m.hr][tx + 4] = h_sa;
    if (ty == 0) {
      Lxx[bf][m.xr][m.xc] = x_xx;
      Lxy[bf][m.xr + 4][m.xc] = x_xy;
      Lxz[bf][m.xr][m.xc] = x_xz;
      if (P) Lsa[bf][m.xr + 4][m.xc] = x_sa;
    }
    __syncthreads();
    // prefetch the next plane and the new queue entries
    float n_zz = 0.f, n_xz = 0.f, n_yz = 0.f, n_sa = 0.f;
    if (k + 1 < k1) {
      load_plane(k + 1, c_xx, c_yy, c_xy, h_yy, h_xy, h_yz, h_sa, x_xx, x_xy, x_xz, x_sa);
      if (m.ld) {
        n_zz = szz[cb + (long)(k + 5) * sz];
        n_xz = sxz[cb + (long)(k + 4) * sz];
        n_yz = syz[cb + (long)(k + 4) * sz];
        if (P) n_sa = sau[cb + (long)(k + 5) * sz];
      }
    }
    if (m.comp) {
      const long id = cb + (long)k * sz;
      const Coef4 cz = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gz + k, gnz, R)) : c4;
      const float dtr = sign * fdtr[id];
      const float a = dtr / dx, bb = dtr / dy, c = dtr / dz;
      const int X = tx + 4, Y = ty + 4;
      float sxx_x = 0.f;
      sxx_x += cx.c0 * (Lxx[bf][ty][X + 1] - Lxx[bf][ty][X]);
      sxx_x += cx.c1 * (Lxx[bf][ty][X + 2] - Lxx[bf][ty][X - 1]);
      sxx_x += cx.c2 * (Lxx[bf][ty][X + 3] - Lxx[bf][ty][X - 2]);
      sxx_x += cx.c3 * (Lxx[bf][ty][X + 4] - Lxx[bf][ty][X - 3]);
      float syy_y = 0.f;
      syy_y += cy.c0 * (Lyy[bf][Y + 1][tx] - Lyy[bf][Y][tx]);
      syy_y += cy.c1 * (Lyy[bf][Y + 2][tx] - Lyy[bf][Y - 1][tx]);
      syy_y += cy.c2 * (Lyy[bf][Y + 3][tx] - Lyy[bf][Y - 2][tx]);
      syy_y += cy.c3 * (Lyy[bf][Y + 4][tx] - Lyy[bf][Y - 3][tx]);
      float szz_z = 0.f;
      szz_z += cz.c0 * (qzz[4] - qzz[3]);
      szz_z += cz.c1 * (qzz[5] - qzz[2]);
      szz_z += cz.c2 * (qzz[6] - qzz[1]);
      szz_z += cz.c3 * (qzz[7] - qzz[0]);
      float syx_x = 0.f;
      syx_x += cx.c0 * (Lxy[bf][Y][X] - Lxy[bf][Y][X - 1]);
      syx_x += cx.c1 * (Lxy[bf][Y][X + 1] - Lxy[bf][Y][X - 2]);
      syx_x += cx.c2 * (Lxy[bf][Y][X + 2] - Lxy[bf][Y][X - 3]);
      syx_x += cx.c3 * (Lxy[bf][Y][X + 3] - Lxy[bf][Y][X - 4]);
      float sxy_y = 0.f;
      sxy_y += cy.c0 * (Lxy[bf][Y][X] - Lxy[bf][Y - 1][X]);
      sxy_y += cy.c1 * (Lxy[bf][Y + 1][X] - Lxy[bf][Y - 2][X]);
      sxy_y += cy.c2 * (Lxy[bf][Y + 2][X] - Lxy[bf][Y - 3][X]);
      sxy_y += cy.c3 * (Lxy[bf][Y + 3][X] - Lxy[bf][Y - 4][X]);
      float sxz_z = 0.f;
      sxz_z += cz.c0 * (qxz[4] - qxz[3]);
      sxz_z += cz.c1 * (qxz[5] - qxz[2]);
      sxz_z += cz.c2 * (qxz[6] - qxz[1]);
      sxz_z += cz.c3 * (qxz[7] - qxz[0]);
      float szx_x = 0.f;
      szx_x += cx.c0 * (Lxz[bf][ty][X] - Lxz[bf][ty][X - 1]);
      szx_x += cx.c1 * (Lxz[bf][ty][X + 1] - Lxz[bf][ty][X - 2]);
      szx_x += cx.c2 * (Lxz[bf][ty][X + 2] - Lxz[bf][ty][X - 3]);
      szx_x += cx.c3 * (Lxz[bf][ty][X + 3] - Lxz[bf][ty][X - 4]);
      float szy_y = 0.f;
      szy_y += cy.c0 * (Lyz[bf][Y][tx] - Lyz[bf][Y - 1][tx]);
      szy_y += cy.c1 * (Lyz[bf][Y + 1][tx] - Lyz[bf][Y - 2][tx]);
      szy_y += cy.c2 * (Lyz[bf][Y + 2][tx] - Lyz[bf][Y - 3][tx]);
      szy_y += cy.c3 * (Lyz[bf][Y + 3][tx] - Lyz[bf][Y - 4][tx]);
      float syz_z = 0.f;
      syz_z += cz.c0 * (qyz[4] - qyz[3]);
      syz_z += cz.c1 * (qyz[5] - qyz[2]);
      syz_z += cz.c2 * (qyz[6] - qyz[1]);
      syz_z += cz.c3 * (qyz[7] - qyz[0]);
      fvx[id] += a * sxx_x + bb * sxy_y + c * sxz_z;
      fvy[id] += a * syx_x + bb * syy_y + c * syz_z;
      fvz[id] += a * szx_x + bb * szy_y + c * szz_z;
      if (P) {
        float sx = 0.f;
        sx += cx.c0 * (Lsa[bf][Y][X + 1] - Lsa[bf][Y][X]);
        sx += cx.c1 * (Lsa[bf][Y][X + 2] - Lsa[bf][Y][X - 1]);
        sx += cx.c2 * (Lsa[bf][Y][X + 3] - Lsa[bf][Y][X - 2]);
        sx += cx.c3 * (Lsa[bf][Y][X + 4] - Lsa[bf][Y][X - 3]);
        float sy_ = 0.f;
        sy_ += cy.c0 * (Lsa[bf][Y + 1][X] - Lsa[bf][Y][X]);
        sy_ += cy.c1 * (Lsa[bf][Y + 2][X] - Lsa[bf][Y - 1][X]);
        sy_ += cy.c2 * (Lsa[bf][Y + 3][X] - Lsa[bf][Y - 2][X]);
        sy_ += cy.c3 * (Lsa[bf][Y + 4][X] - Lsa[bf][Y - 3][X]);
        float sz_ = 0.f;
        sz_ += cz.c0 * (qsa[4] - qsa[3]);
        sz_ += cz.c1 * (qsa[5] - qsa[2]);
        sz_ += cz.c2 * (qsa[6] - qsa[1]);
        sz_ += cz.c3 * (qsa[7] - qsa[0]);
        fvxp[id] += a * sx;
        fvyp[id] += bb * sy_;
        fvzp[id] += c * sz_;
      }
    }
#pragma unroll
    for (int r = 0; r < 7; r++) {
      qzz[r] = qzz[r + 1];
      qxz[r] = qxz[r + 1];
      qyz[r] = qyz[r + 1];
      if (P) qsa[r] = qsa[r + 1];
    }
    qzz[7] = n_zz;
    qxz[7] = n_xz;
    qyz[7] = n_yz;
    if (P) qsa[7] = n_sa;
  }
}

template <int MODE>
__global__ __launch_bounds__(LX *LY) void k_stress_lds(const PhyDesc *__restrict__ descs,
                                                       const int *__restrict__ starts, int n, float dx, float dy,
                                                       float dz, float dt, int gnx, int gny, int gnz, int R,
                                                       int kz) {
  __shared__ float Lx[2][HYW][HXW];
  __shared__ float Ly[2][HYW][HXW];
  __shared__ float Lz[2][HYW][HXW];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const PhyDesc &D = descs[find_desc(starts, n, b)];
  gnx = D.gnx; gny = D.gny; gnz = D.gnz; R = D.R;  // per descriptor: fused batches mix grids
  const int t = b - D.blk0;
  const int tiles_x = (D.nx + LX - 1) / LX, tiles_y = (D.ny + LY - 1) / LY, per = tiles_x * tiles_y;
  const int zc = t / per, rem = t - zc * per, tyi = rem / tiles_x, txi = rem - tyi * tiles_x;
  const int x0 = txi * LX, y0 = tyi * LY, k0 = zc * kz, k1 = min(k0 + kz, D.nz);
  const LdsMap m = lds_map(D, x0, y0);
  const int tx = threadIdx.x % LX, ty = threadIdx.x / LX;
  const long sy = D.sy, sz = D.sz;
  const gfloat *vx = G(D.f[VX]), *vy = G(D.f[VY]), *vz = G(D.f[VZ]);
  const long cb = (long)(D.ox + m.i) + sy * (D.oy + m.j) + sz * D.oz;
  const long hb = (long)(D.ox + m.i) + sy * (D.oy + m.hrow) + sz * D.oz;
  const long xb = (long)(D.ox + m.xcol) + sy * (D.oy + y0 + m.xr) + sz * D.oz;
  const float a = dt / dx, bb = dt / dy, c = dt / dz;
  const Coef4 c4 = coef_row(4);
  const Coef4 cx = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gx + min(m.i, D.nx - 1), gnx, R)) : c4;
  const Coef4 cy = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gy + min(m.j, D.ny - 1), gny, R)) : c4;
  // queues: qx, qy planes k-3..k+4 (forward-staggered in z), qz planes k-4..k+3
  float qx[8], qy[8], qz[8];
#pragma unroll
  for (int r = 0; r < 8; r++) {
    qx[r] = m.ld ? vx[cb + (long)(k0 - 3 + r) * sz] : 0.f;
    qy[r] = m.ld ? vy[cb + (long)(k0 - 3 + r) * sz] : 0.f;
    qz[r] = m.ld ? vz[cb + (long)(k0 - 4 + r) * sz] : 0.f;
  }
  float h_x = 0.f, h_y = 0.f, h_z = 0.f, e_x = 0.f, e_y = 0.f, e_z = 0.f;
  auto load_halo = [&](int k) {
    const long kk = (long)k * sz;
    if (m.hld) { h_x = vx[hb + kk]; h_y = vy[hb + kk]; h_z = vz[hb + kk]; }
    if (m.xld) { e_x = vx[xb + kk]; e_y = vy[xb + kk]; e_z = vz[xb + kk]; }
  };
  load_halo(k0);
  gfloat *fsxx = G(D.f[SXX]), *fsyy = G(D.f[SYY]), *fszz = G(D.f[SZZ]);
  gfloat *fsxy = G(D.f[SXY]), *fsxz = G(D.f[SXZ]), *fsyz = G(D.f[SYZ]), *fsau = G(D.f[SAU]);
  const gfloat *flam = G(D.lam), *fmu = G(D.mu);
  for (int k = k0; k < k1; k++) {
    const int bf = (k - k0) & 1;
    Lx[bf][ty + 4][tx + 4] = qx[3];
    Ly[bf][ty + 4][tx + 4] = qy[3];
    Lz[bf][ty + 4][tx + 4] = qz[4];
    Lx[bf][m.hr][tx + 4] = h_x;
    Ly[bf][m.hr][tx + 4] = h_y;
    Lz[bf][m.hr][tx + 4] = h_z;
    if (ty == 0) {
      Lx[bf][m.xr + 4][m.xc] = e_x;
      Ly[bf][m.xr + 4][m.xc] = e_y;
      Lz[bf][m.xr + 4][m.xc] = e_z;
    }
    __syncthreads();
    float n_x = 0.f, n_y = 0.f, n_z = 0.f;
    if (k + 1 < k1) {
      load_halo(k + 1);
      if (m.ld) {
        n_x = vx[cb + (long)(k + 5) * sz];
        n_y = vy[cb + (long)(k + 5) * sz];
        n_z = vz[cb + (long)(k + 4) * sz];
      }
    }
    if (m.comp) {
      const long id = cb + (long)k * sz;
      const Coef4 cz = MODE == RTM_BACKWARD ? coef_row(vowr_order(D.gz + k, gnz, R)) : c4;
      const float lm = flam[id], mu = fmu[id];
      const float l2m = lm + 2.0f * mu;
      const int X = tx + 4, Y = ty + 4;
      // vx: x arm backward-staggered, y and z arms forward-staggered
      float vx_x = 0.f;
      vx_x += cx.c0 * (Lx[bf][Y][X] - Lx[bf][Y][X - 1]);
      vx_x += cx.c1 * (Lx[bf][Y][X + 1] - Lx[bf][Y][X - 2]);
      vx_x += cx.c2 * (Lx[bf][Y][X + 2] - Lx[bf][Y][X - 3]);
      vx_x += cx.c3 * (Lx[bf][Y][X + 3] - Lx[bf][Y][X - 4]);
      float vx_y = 0.f;
      vx_y += cy.c0 * (Lx[bf][Y + 1][X] - Lx[bf][Y][X]);
      vx_y += cy.c1 * (Lx[bf][Y + 2][X] - Lx[bf][Y - 1][X]);
      vx_y += cy.c2 * (Lx[bf][Y + 3][X] - Lx[bf][Y - 2][X]);
      vx_y += cy.c3 * (Lx[bf][Y + 4][X] - Lx[bf][Y - 3][X]);
      float vx_z = 0.f;
      vx_z += cz.c0 * (qx[4] - qx[3]);
      vx_z += cz.c1 * (qx[5] - qx[2]);
      vx_z += cz.c2 * (qx[6] - qx[1]);
      vx_z += cz.c3 * (qx[7] - qx[0]);
      // vy: x and z forward, y backward
      float vy_x = 0.f;
      vy_x += cx.c0 * (Ly[bf][Y][X + 1] - Ly[bf][Y][X]);
      vy_x += cx.c1 * (Ly[bf][Y][X + 2] - Ly[bf][Y][X - 1]);
      vy_x += cx.c2 * (Ly[bf][Y][X + 3] - Ly[bf][Y][X - 2]);
      vy_x += cx.c3 * (Ly[bf][Y][X + 4] - Ly[bf][Y][X - 3]);
      float vy_y = 0.f;
      vy_y += cy.c0 * (Ly[bf][Y][X] - Ly[bf][Y - 1][X]);
      vy_y += cy.c1 * (Ly[bf][Y + 1][X] - Ly[bf][Y - 2][X]);
      vy_y += cy.c2 * (Ly[bf][Y + 2][X] - Ly[bf][Y - 3][X]);
      vy_y += cy.c3 * (Ly[bf][Y + 3][X] - Ly[bf][Y - 4][X]);
      float vy_z = 0.f;
      vy_z += cz.c0 * (qy[4] - qy[3]);
      vy_z += cz.c1 * (qy[5] - qy[2]);
      vy_z += cz.c2 * (qy[6] - qy[1]);
      vy_z += cz.c3 * (qy[7] - qy[0]);
      // vz: x and y forward, z backward
      float vz_x = 0.f;
      vz_x += cx.c0 * (Lz[bf][Y][X + 1] - Lz[bf][Y][X]);
      vz_x += cx.c1 * (Lz[bf][Y][X + 2] - Lz[bf][Y][X - 1]);
      vz_x += cx.c2 * (Lz[bf][Y][X + 3] - Lz[bf][Y][X - 2]);
      vz_x += cx.c3 * (Lz[bf][Y][X + 4] - Lz[bf][Y][X - 3]);
      float vz_y = 0.f;
      vz_y += cy.c0 * (Lz[bf][Y + 1][X] - Lz[bf][Y][X]);
      vz_y += cy.c1 * (Lz[bf][Y + 2][X] - Lz[bf][Y - 1][X]);
      vz_y += cy.c2 * (Lz[bf][Y + 3][X] - Lz[bf][Y - 2][X]);
      vz_y += cy.c3 * (Lz[bf][Y + 4][X] - Lz[bf][Y - 3][X]);
      float vz_z = 0.f;
      vz_z += cz.c0 * (qz[4] - qz[3]);
      vz_z += cz.c1 * (qz[5] - qz[2]);
      vz_z += cz.c2 * (qz[6] - qz[1]);
      vz_z += cz.c3 * (qz[7] - qz[0]);
      fsxx[id] += a * l2m * vx_x + bb * lm * vy_y + c * lm * vz_z;
      fsyy[id] += a * lm * vx_x + bb * l2m * vy_y + c * lm * vz_z;
      fszz[id] += a * lm * vx_x + bb * lm * vy_y + c * l2m * vz_z;
      fsxy[id] += a * mu * vy_x + bb * mu * vx_y;
      fsxz[id] += a * mu * vz_x + c * mu * vx_z;
      fsyz[id] += bb * mu * vz_y + c * mu * vy_z;
      if (rtm_mode(MODE)) fsau[id] += a * l2m * vx_x + bb * l2m * vy_y + c * l2m * vz_z;
    }
#pragma unroll
    for (int r = 0; r < 7; r++) {
      qx[r] = qx[r + 1];
      qy[r] = qy[r + 1];
      qz[r] = qz[r + 1];
    }
    qx[7] = n_x;
    qy[7] = n_y;
    qz[7] = n_z;
  }
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
// Work split: a thread owns one (x, y) column of a slab and marches CZ
// consecutive planes of it; blocks cover kBlock consecutive cells of a slab
// plane (x fastest, so every wavefront reads contiguous x runs) times one z
// chunk.  The z taps of the 4th-order stencil come from register queues of
// CZ + 3 planes loaded once per column; x and y taps are read through the
// caches.  CZ = 1 is one cell per thread.  Slabs hold fewer than 2^31 cells (checked on the
// host), so the index math is 32-bit.
//
// Which slab, z chunk and plane tile a workgroup covers: with ord.pieces,
// the XCD-partitioned order (PmlOrder); else block b of the table's own
// order after the group-interleaved remap grp.  Returns false for the
// surplus workgroups of an XCD with fewer blocks than the largest.
__device__ __forceinline__ bool pml_block(const SlabDesc *__restrict__ descs, const int *__restrict__ starts,
                                          int n, int grp, const PmlOrder &ord, int cz, int &di, int &zc,
                                          int &tile) {
  if (ord.pieces) {
    const int x = (int)blockIdx.x & 7, local = (int)blockIdx.x >> 3;
    if (local >= ord.xn[x]) return false;
    int lo = ord.xoff[x], hi = ord.xoff[x + 1] - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (ord.pieces[mid].lstart <= local) lo = mid; else hi = mid - 1;
    }
    const PmlPiece P = ord.pieces[lo];
    const int off = local - P.lstart;
    const int q = off / P.tcnt;
    di = P.desc;
    zc = P.p0 + q;
    tile = P.t0 + (off - q * P.tcnt);
    return true;
  }
  const int b = xcd_group_remap(blockIdx.x, gridDim.x, grp);
  di = find_desc(starts, n, b);
  const SlabDesc &D = descs[di];
  const int tiles = pml_plane_tiles(D.ni, D.nj, cz);
  const int t = b - D.blk0;
  zc = t / tiles;
  tile = t - zc * tiles;
  return true;
}

template <int CZ>
__device__ __forceinline__ bool slab_column(const SlabDesc &D, int zc, int tile, int &ci, int &ei, int &ej,
                                            int &kc0, int &kend) {
  const int plane = D.ni * D.nj;
  ci = tile * kBlock + (int)threadIdx.x;
  const int r = ci / D.ni;
  ei = D.i0 + (ci - r * D.ni);
  ej = D.j0 + r;
  kc0 = zc * CZ;
  kend = min(kc0 + CZ, D.nk);
  return ci < plane;
}

// zq[r] = f(slab plane kc0 + lo + r) for the planes the outputs need
// (up to kend + 1); the rest read as 0 and are never used.
template <int CZ, typename GT>
__device__ __forceinline__ void load_pml_zq(float (&zq)[CZ + 3], const GT *__restrict__ f, long base,
                                            long sz, int kc0, int lo, int kend) {
#pragma unroll
  for (int r = 0; r < CZ + 3; r++) zq[r] = (kc0 + lo + r <= kend + 1) ? ld(f, base + (long)(lo + r) * sz) : 0.f;
}

// Extra attributes of the CPML kernels (occupancy experiments), empty by default.
#ifndef DDX_PML_ATTR
#define DDX_PML_ATTR
#endif

// Memory variables and CPML coefficients of one damped axis of a cell.  The
// kernels read them, and the updated fields, before any store: on this
// target the memory counter also tracks stores, so a read issued after a
// store waits for it.
template <typename GT>
struct PsiCell {
  GT *p;
  float a, b, ah, bh, m0, m1, m2;
};
template <typename GT>
__device__ __forceinline__ PsiCell<GT> psi_cell(GT *p, const gfloat *q, long ncell) {
  return {p, q[0], q[ncell], q[2 * ncell], q[3 * ncell], ld(p, 0), ld(p, ncell), ld(p, 2 * ncell)};
}

// S: storage format of fields and memory variables; rv: velocity-over-stress
// scale ratio (1 for F32), folded into dt/rho.
template <int CZ, class S>
__global__ __launch_bounds__(kBlock) DDX_PML_ATTR void k_vel_pml(const SlabDesc *__restrict__ descs,
                                                    const int *__restrict__ starts, int n, float dx, float dy,
                                                    float dz, float rv, int grp, PmlOrder ord) {
  typedef typename GPtr<S>::T gT;
  int di, zc, tile;
  if (!pml_block(descs, starts, n, grp, ord, CZ, di, zc, tile)) return;
  const SlabDesc &D = descs[di];
  int ci, ei, ej, kc0, kend;
  if (!slab_column<CZ>(D, zc, tile, ci, ei, ej, kc0, kend)) return;
  const long ncell = (long)D.ni * D.nj * D.nk;
  const int plane = D.ni * D.nj;
  const long sy = D.sy, sz = D.sz;
  const long base = (long)ei + sy * ej + sz * (D.k0 + kc0);
  const float c0 = c_fd[2][0], c1 = c_fd[2][1];
  const gT *sxx = (const gT *)D.f[SXX], *syy = (const gT *)D.f[SYY], *szz = (const gT *)D.f[SZZ];
  const gT *sxy = (const gT *)D.f[SXY], *sxz = (const gT *)D.f[SXZ], *syz = (const gT *)D.f[SYZ];
  gT *fvx = (gT *)D.f[VX], *fvy = (gT *)D.f[VY], *fvz = (gT *)D.f[VZ];
  gT *psi = (gT *)D.psi;
  const gfloat *fdtr = G(D.dtr);
  float qzz[CZ + 3], qxz[CZ + 3], qyz[CZ + 3];
  load_pml_zq<CZ>(qzz, szz, base, sz, kc0, -1, kend);  // szz: forward-staggered, planes k-1 .. k+2
  load_pml_zq<CZ>(qxz, sxz, base, sz, kc0, -2, kend);  // sxz, syz: backward-staggered, planes k-2 .. k+1
  load_pml_zq<CZ>(qyz, syz, base, sz, kc0, -2, kend);
  const bool inx = (D.axes & 1) && (ei < D.xlo || ei >= D.xhi);
  const bool iny = (D.axes & 2) && (ej < D.ylo || ej >= D.yhi);
  const int sx_slot = 0, sy_slot = (D.axes & 1) ? 1 : 0, sz_slot = sy_slot + ((D.axes & 2) ? 1 : 0);
  // Two phases: every plane's loads and results first, then every store,
  // so that no load of a later plane waits behind an earlier plane's
  // stores (the memory counter tracks both).  A memory variable's new value
  // enters the sums as the value stored (quantized to the storage format).
  float nv[CZ][3], np[CZ][9];
  bool inz[CZ];
#pragma unroll
  for (int kk = 0; kk < CZ; kk++) {
    inz[kk] = false;
    if (kc0 + kk >= kend) break;
    const int ek = D.k0 + kc0 + kk;
    const long id = base + (long)kk * sz;
    const long c = (long)ci + (long)plane * (kc0 + kk);
    const float ovx = ld(fvx, id), ovy = ld(fvy, id), ovz = ld(fvz, id);
    const float dtr = std::is_same<S, F32>::value ? fdtr[id] : rv * fdtr[id];
    float sxx_x = dp2(sxx, id, 1, c0, c1), syx_x = dm2(sxy, id, 1, c0, c1), szx_x = dm2(sxz, id, 1, c0, c1);
    float syy_y = dp2(syy, id, sy, c0, c1), sxy_y = dm2(sxy, id, sy, c0, c1), szy_y = dm2(syz, id, sy, c0, c1);
    float szz_z = 0.f, sxz_z = 0.f, syz_z = 0.f;
    szz_z += c0 * (qzz[kk + 2] - qzz[kk + 1]);
    szz_z += c1 * (qzz[kk + 3] - qzz[kk]);
    sxz_z += c0 * (qxz[kk + 2] - qxz[kk + 1]);
    sxz_z += c1 * (qxz[kk + 3] - qxz[kk]);
    syz_z += c0 * (qyz[kk + 2] - qyz[kk + 1]);
    syz_z += c1 * (qyz[kk + 3] - qyz[kk]);
    inz[kk] = (D.axes & 4) && (ek < D.zlo || ek >= D.zhi);
    PsiCell<gT> mx{}, my{}, mz{};
    if (inx) mx = psi_cell(psi + (long)(sx_slot * 6) * ncell + c, G(D.coef) + (long)(sx_slot * 4) * ncell + c, ncell);
    if (iny) my = psi_cell(psi + (long)(sy_slot * 6) * ncell + c, G(D.coef) + (long)(sy_slot * 4) * ncell + c, ncell);
    if (inz[kk])
      mz = psi_cell(psi + (long)(sz_slot * 6) * ncell + c, G(D.coef) + (long)(sz_slot * 4) * ncell + c, ncell);
    sxx_x = sxx_x / dx; syy_y = syy_y / dy; szz_z = szz_z / dz;
    sxy_y = sxy_y / dy; sxz_z = sxz_z / dz; syx_x = syx_x / dx;
    syz_z = syz_z / dz; szx_x = szx_x / dx; szy_y = szy_y / dy;
    if (inx) {
      const float p0 = quant(fvx, mx.bh * mx.m0 + mx.ah * sxx_x);
      const float p1 = quant(fvx, mx.b * mx.m1 + mx.a * syx_x);
      const float p2 = quant(fvx, mx.b * mx.m2 + mx.a * szx_x);
      np[kk][0] = p0; np[kk][1] = p1; np[kk][2] = p2;
      sxx_x += p0; syx_x += p1; szx_x += p2;
    }
    if (iny) {
      const float p0 = quant(fvx, my.b * my.m0 + my.a * sxy_y);
      const float p1 = quant(fvx, my.bh * my.m1 + my.ah * syy_y);
      const float p2 = quant(fvx, my.b * my.m2 + my.a * szy_y);
      np[kk][3] = p0; np[kk][4] = p1; np[kk][5] = p2;
      sxy_y += p0; syy_y += p1; szy_y += p2;
    }
    if (inz[kk]) {
      const float p0 = quant(fvx, mz.b * mz.m0 + mz.a * sxz_z);
      const float p1 = quant(fvx, mz.b * mz.m1 + mz.a * syz_z);
      const float p2 = quant(fvx, mz.bh * mz.m2 + mz.ah * szz_z);
      np[kk][6] = p0; np[kk][7] = p1; np[kk][8] = p2;
      sxz_z += p0; syz_z += p1; szz_z += p2;
    }
    nv[kk][0] = ovx + dtr * (sxx_x + sxy_y + sxz_z);
    nv[kk][1] = ovy + dtr * (syx_x + syy_y + syz_z);
    nv[kk][2] = ovz + dtr * (szx_x + szy_y + szz_z);
  }
#pragma unroll
  for (int kk = 0; kk < CZ; kk++) {
    if (kc0 + kk >= kend) break;
    const int ek = D.k0 + kc0 + kk;
    const long id = base + (long)kk * sz;
    const long c = (long)ci + (long)plane * (kc0 + kk);
    if (inx) {
      gT *q = psi + (long)(sx_slot * 6) * ncell + c;
      stq(q, 0, np[kk][0]); stq(q, ncell, np[kk][1]); stq(q, 2 * ncell, np[kk][2]);
    }
    if (iny) {
      gT *q = psi + (long)(sy_slot * 6) * ncell + c;
      stq(q, 0, np[kk][3]); stq(q, ncell, np[kk][4]); stq(q, 2 * ncell, np[kk][5]);
    }
    if (inz[kk]) {
      gT *q = psi + (long)(sz_slot * 6) * ncell + c;
      stq(q, 0, np[kk][6]); stq(q, ncell, np[kk][7]); stq(q, 2 * ncell, np[kk][8]);
    }
    stq(fvx, id, nv[kk][0]);
    stq(fvy, id, nv[kk][1]);
    stq(fvz, id, nv[kk][2]);
    if (D.mir) {
      const MirCell m = mir_cell(D.mir, ei, ej, ek);
      if (m.any) { mir_put<gT>(m, VX, nv[kk][0]); mir_put<gT>(m, VY, nv[kk][1]); mir_put<gT>(m, VZ, nv[kk][2]); }
    }
  }
}

// `dt` carries the stress-over-velocity scale ratio of the 16-bit formats.
template <int WITH_SAU, int CZ, class S>
__global__ __launch_bounds__(kBlock) DDX_PML_ATTR void k_stress_pml(const SlabDesc *__restrict__ descs,
                                                       const int *__restrict__ starts, int n, float dx,
                                                       float dy, float dz, float dt, int grp, PmlOrder ord) {
  typedef typename GPtr<S>::T gT;
  int di, zc, tile;
  if (!pml_block(descs, starts, n, grp, ord, CZ, di, zc, tile)) return;
  const SlabDesc &D = descs[di];
  int ci, ei, ej, kc0, kend;
  if (!slab_column<CZ>(D, zc, tile, ci, ei, ej, kc0, kend)) return;
  const long ncell = (long)D.ni * D.nj * D.nk;
  const int plane = D.ni * D.nj;
  const long sy = D.sy, sz = D.sz;
  const long base = (long)ei + sy * ej + sz * (D.k0 + kc0);
  const float c0 = c_fd[2][0], c1 = c_fd[2][1];
  const gT *vx = (const gT *)D.f[VX], *vy = (const gT *)D.f[VY], *vz = (const gT *)D.f[VZ];
  gT *fsxx = (gT *)D.f[SXX], *fsyy = (gT *)D.f[SYY], *fszz = (gT *)D.f[SZZ];
  gT *fsxy = (gT *)D.f[SXY], *fsxz = (gT *)D.f[SXZ], *fsyz = (gT *)D.f[SYZ], *fsau = (gT *)D.f[SAU];
  gT *psi = (gT *)D.psi;
  const gfloat *flam = G(D.lam), *fmu = G(D.mu);
  float qx[CZ + 3], qy[CZ + 3], qz[CZ + 3];
  load_pml_zq<CZ>(qx, vx, base, sz, kc0, -1, kend);  // vx, vy: forward-staggered in z
  load_pml_zq<CZ>(qy, vy, base, sz, kc0, -1, kend);
  load_pml_zq<CZ>(qz, vz, base, sz, kc0, -2, kend);  // vz: backward-staggered in z
  const bool inx = (D.axes & 1) && (ei < D.xlo || ei >= D.xhi);
  const bool iny = (D.axes & 2) && (ej < D.ylo || ej >= D.yhi);
  const int sx_slot = 0, sy_slot = (D.axes & 1) ? 1 : 0, sz_slot = sy_slot + ((D.axes & 2) ? 1 : 0);
  // loads and results of every plane first, then the stores (as k_vel_pml)
  float ns[CZ][7], np[CZ][9];
  bool inz[CZ];
#pragma unroll
  for (int kk = 0; kk < CZ; kk++) {
    inz[kk] = false;
    if (kc0 + kk >= kend) break;
    const int ek = D.k0 + kc0 + kk;
    const long id = base + (long)kk * sz;
    const long c = (long)ci + (long)plane * (kc0 + kk);
    float vx_x = dm2(vx, id, 1, c0, c1), vy_x = dp2(vy, id, 1, c0, c1), vz_x = dp2(vz, id, 1, c0, c1);
    float vy_y = dm2(vy, id, sy, c0, c1), vx_y = dp2(vx, id, sy, c0, c1), vz_y = dp2(vz, id, sy, c0, c1);
    float vz_z = 0.f, vx_z = 0.f, vy_z = 0.f;
    vz_z += c0 * (qz[kk + 2] - qz[kk + 1]);
    vz_z += c1 * (qz[kk + 3] - qz[kk]);
    vx_z += c0 * (qx[kk + 2] - qx[kk + 1]);
    vx_z += c1 * (qx[kk + 3] - qx[kk]);
    vy_z += c0 * (qy[kk + 2] - qy[kk + 1]);
    vy_z += c1 * (qy[kk + 3] - qy[kk]);
    const float lm = flam[id], mu = fmu[id];
    const float l2m = lm + 2.0f * mu;
    const float oxx = ld(fsxx, id), oyy = ld(fsyy, id), ozz = ld(fszz, id);
    const float oxy = ld(fsxy, id), oxz = ld(fsxz, id), oyz = ld(fsyz, id);
    const float oau = WITH_SAU ? ld(fsau, id) : 0.f;
    inz[kk] = (D.axes & 4) && (ek < D.zlo || ek >= D.zhi);
    PsiCell<gT> mx{}, my{}, mz{};
    if (inx)
      mx = psi_cell(psi + (long)(sx_slot * 6 + 3) * ncell + c, G(D.coef) + (long)(sx_slot * 4) * ncell + c, ncell);
    if (iny)
      my = psi_cell(psi + (long)(sy_slot * 6 + 3) * ncell + c, G(D.coef) + (long)(sy_slot * 4) * ncell + c, ncell);
    if (inz[kk])
      mz = psi_cell(psi + (long)(sz_slot * 6 + 3) * ncell + c, G(D.coef) + (long)(sz_slot * 4) * ncell + c, ncell);
    vx_x = vx_x / dx; vy_y = vy_y / dy; vz_z = vz_z / dz;
    vy_x = vy_x / dx; vz_x = vz_x / dx; vx_y = vx_y / dy;
    vz_y = vz_y / dy; vx_z = vx_z / dz; vy_z = vy_z / dz;
    if (inx) {
      const float p0 = quant(fsxx, mx.b * mx.m0 + mx.a * vx_x);
      const float p1 = quant(fsxx, mx.bh * mx.m1 + mx.ah * vy_x);
      const float p2 = quant(fsxx, mx.bh * mx.m2 + mx.ah * vz_x);
      np[kk][0] = p0; np[kk][1] = p1; np[kk][2] = p2;
      vx_x += p0; vy_x += p1; vz_x += p2;
    }
    if (iny) {
      const float p0 = quant(fsxx, my.bh * my.m0 + my.ah * vx_y);
      const float p1 = quant(fsxx, my.b * my.m1 + my.a * vy_y);
      const float p2 = quant(fsxx, my.bh * my.m2 + my.ah * vz_y);
      np[kk][3] = p0; np[kk][4] = p1; np[kk][5] = p2;
      vx_y += p0; vy_y += p1; vz_y += p2;
    }
    if (inz[kk]) {
      const float p0 = quant(fsxx, mz.bh * mz.m0 + mz.ah * vx_z);
      const float p1 = quant(fsxx, mz.bh * mz.m1 + mz.ah * vy_z);
      const float p2 = quant(fsxx, mz.b * mz.m2 + mz.a * vz_z);
      np[kk][6] = p0; np[kk][7] = p1; np[kk][8] = p2;
      vx_z += p0; vy_z += p1; vz_z += p2;
    }
    ns[kk][0] = oxx + dt * (l2m * vx_x + lm * vy_y + lm * vz_z);
    ns[kk][1] = oyy + dt * (lm * vx_x + l2m * vy_y + lm * vz_z);
    ns[kk][2] = ozz + dt * (lm * vx_x + lm * vy_y + l2m * vz_z);
    ns[kk][3] = WITH_SAU ? oau + dt * (l2m * vx_x + l2m * vy_y + l2m * vz_z) : 0.f;
    ns[kk][4] = oxy + dt * mu * (vy_x + vx_y);
    ns[kk][5] = oxz + dt * mu * (vz_x + vx_z);
    ns[kk][6] = oyz + dt * mu * (vz_y + vy_z);
  }
#pragma unroll
  for (int kk = 0; kk < CZ; kk++) {
    if (kc0 + kk >= kend) break;
    const int ek = D.k0 + kc0 + kk;
    const long id = base + (long)kk * sz;
    const long c = (long)ci + (long)plane * (kc0 + kk);
    if (inx) {
      gT *q = psi + (long)(sx_slot * 6 + 3) * ncell + c;
      stq(q, 0, np[kk][0]); stq(q, ncell, np[kk][1]); stq(q, 2 * ncell, np[kk][2]);
    }
    if (iny) {
      gT *q = psi + (long)(sy_slot * 6 + 3) * ncell + c;
      stq(q, 0, np[kk][3]); stq(q, ncell, np[kk][4]); stq(q, 2 * ncell, np[kk][5]);
    }
    if (inz[kk]) {
      gT *q = psi + (long)(sz_slot * 6 + 3) * ncell + c;
      stq(q, 0, np[kk][6]); stq(q, ncell, np[kk][7]); stq(q, 2 * ncell, np[kk][8]);
    }
    stq(fsxx, id, ns[kk][0]);
    stq(fsyy, id, ns[kk][1]);
    stq(fszz, id, ns[kk][2]);
    if (WITH_SAU) stq(fsau, id, ns[kk][3]);
    stq(fsxy, id, ns[kk][4]);
    stq(fsxz, id, ns[kk][5]);
    stq(fsyz, id, ns[kk][6]);
    if (D.mir) {
      const MirCell m = mir_cell(D.mir, ei, ej, ek);
      if (m.any) {
        mir_put<gT>(m, SXX, ns[kk][0]); mir_put<gT>(m, SYY, ns[kk][1]); mir_put<gT>(m, SZZ, ns[kk][2]);
        if (WITH_SAU) mir_put<gT>(m, SAU, ns[kk][3]);
        mir_put<gT>(m, SXY, ns[kk][4]); mir_put<gT>(m, SXZ, ns[kk][5]); mir_put<gT>(m, SYZ, ns[kk][6]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// CPML halves on LDS-staged plane tiles (FP32 storage; work split
// kPmlTiled + P).  A block owns a tile of TX x TY cells of a slab plane
// (TX = pml_tile_x(ni), TY = kBlock / TX) and marches P planes of it.  Per
// plane, the fields whose x or y derivative the update takes are staged in
// LDS with the halo of the 4th-order taps (two cells on the side the stagger
// points away from, one on the other); the z arm comes from per-thread
// register queues, whose centre values also fill the tiles of the fields
// that have them.  The loads of plane k + 1 (halo cells, queue heads and the
// cell's own streams) are issued before the stores of plane k, so that no
// load waits behind a store; the tiles are double-buffered, one barrier per
// plane.  Operands and order of operations are those of k_vel_pml and
// k_stress_pml, so the results are bit-identical to theirs.
// ---------------------------------------------------------------------------
constexpr int kPmlLds = 544;  // (TX + 4) (TY + 4) floats at most, TX in [4, 64], TX TY <= kBlock

// One thread's cell of a tile halo: `lds` its index in the tile (-1: the
// thread holds none of this halo), `off` its offset from the tile origin.
// The halo of an x arm spans the tile's h rows, `lo` cells below column 0
// and `hi` from column w on (w x h: the tile's cells inside the slab); the
// halo of a y arm likewise spans its w columns.  Only cells the stencils
// read are loaded: the same cells as k_vel_pml / k_stress_pml read.
struct TileSlot {
  int lds, off;
};
__device__ __forceinline__ TileSlot tile_slot(int t, bool xarm, int lo, int hi, int w, int h, int px, int sy) {
  TileSlot s{-1, 0};
  const int k = lo + hi;
  int row, col;
  if (xarm) {
    if (t >= k * h) return s;
    row = t / k;
    const int c = t - row * k;
    col = c < lo ? c - lo : w + c - lo;
  } else {
    if (t >= k * w) return s;
    const int r = t / w;
    col = t - r * w;
    row = r < lo ? r - lo : h + r - lo;
  }
  s.lds = (row + 2) * px + col + 2;
  s.off = col + row * sy;
  return s;
}

// 4th-order differences on an LDS tile, as dp2 / dm2.
__device__ __forceinline__ float tp2(const float *s, int i, int st, float c0, float c1) {
  float a = 0.f;
  a += c0 * (s[i + st] - s[i]);
  a += c1 * (s[i + 2 * st] - s[i - st]);
  return a;
}
__device__ __forceinline__ float tm2(const float *s, int i, int st, float c0, float c1) {
  float a = 0.f;
  a += c0 * (s[i] - s[i - st]);
  a += c1 * (s[i + st] - s[i - 2 * st]);
  return a;
}

// A block's tile: slab, plane tile and z chunk, and this thread's place in it.
struct PmlTile {
  int x0, y0, w, h, tx, ty, px, me, kc0, np;
  bool live;
};
__device__ __forceinline__ PmlTile pml_tile(const SlabDesc &D, int zc, int tile, int planes) {
  PmlTile T;
  const int TX = pml_tile_x(D.ni), TY = kBlock / TX;
  const int ntx = (D.ni + TX - 1) / TX;
  const int r = tile / ntx;
  T.x0 = (tile - r * ntx) * TX;
  T.y0 = r * TY;
  T.w = min(TX, D.ni - T.x0);
  T.h = min(TY, D.nj - T.y0);
  const int t = (int)threadIdx.x;
  T.ty = t / TX;
  T.tx = t - T.ty * TX;
  T.live = T.tx < T.w && T.ty < T.h;
  T.px = TX + 4;
  T.me = (T.ty + 2) * T.px + T.tx + 2;
  T.kc0 = zc * planes;
  T.np = min(planes, D.nk - T.kc0);
  return T;
}

// One plane's loads for a cell of the velocity update.
struct VelPlane {
  float sxx, sxy, syy;                   // tile centres (sxz, syz come from the queues)
  float hxx, hxy, hxz, hyxy, hyy, hyz;   // halo cells: sxx (x, 1|2), sxy, sxz (x, 2|1); sxy, syz (y, 2|1), syy (y, 1|2)
  float nzz, nxz, nyz;                   // queue heads: szz plane k + 2, sxz and syz plane k + 1
  float ovx, ovy, ovz, dtr;
  PsiCell<gfloat> mx, my, mz;
  bool inz;
};

template <int P>
__global__ __launch_bounds__(kBlock) DDX_PML_ATTR void k_vel_pml_t(const SlabDesc *__restrict__ descs,
                                                      const int *__restrict__ starts, int n, float dx, float dy,
                                                      float dz, int grp, PmlOrder ord) {
  __shared__ float lt[2][5][kPmlLds];  // sxx, sxy, sxz, syy, syz
  int di, zc, tile;
  if (!pml_block(descs, starts, n, grp, ord, kPmlTiled + P, di, zc, tile)) return;
  const SlabDesc &D = descs[di];
  const PmlTile T = pml_tile(D, zc, tile, P);
  const long ncell = (long)D.ni * D.nj * D.nk;
  const int plane = D.ni * D.nj;
  const int sy = (int)D.sy;
  const long sz = D.sz;
  const long tb0 = (long)(D.i0 + T.x0) + D.sy * (D.j0 + T.y0) + sz * (D.k0 + T.kc0);  // tile origin
  const long id0 = tb0 + T.tx + (long)T.ty * sy;                                       // this cell
  const long c00 = (long)(T.x0 + T.tx) + (long)D.ni * (T.y0 + T.ty) + (long)plane * T.kc0;
  const int ei = D.i0 + T.x0 + T.tx, ej = D.j0 + T.y0 + T.ty;
  const int t = (int)threadIdx.x;
  const TileSlot x12 = tile_slot(t, true, 1, 2, T.w, T.h, T.px, sy);
  const TileSlot x21 = tile_slot(t, true, 2, 1, T.w, T.h, T.px, sy);
  const TileSlot y12 = tile_slot(t, false, 1, 2, T.w, T.h, T.px, sy);
  const TileSlot y21 = tile_slot(t, false, 2, 1, T.w, T.h, T.px, sy);
  const float c0 = c_fd[2][0], c1 = c_fd[2][1];
  const gfloat *sxx = (const gfloat *)D.f[SXX], *syy = (const gfloat *)D.f[SYY], *szz = (const gfloat *)D.f[SZZ];
  const gfloat *sxy = (const gfloat *)D.f[SXY], *sxz = (const gfloat *)D.f[SXZ], *syz = (const gfloat *)D.f[SYZ];
  gfloat *fvx = (gfloat *)D.f[VX], *fvy = (gfloat *)D.f[VY], *fvz = (gfloat *)D.f[VZ];
  gfloat *psi = (gfloat *)D.psi;
  const gfloat *fdtr = G(D.dtr), *coef = G(D.coef);
  const bool inx = (D.axes & 1) && (ei < D.xlo || ei >= D.xhi);
  const bool iny = (D.axes & 2) && (ej < D.ylo || ej >= D.yhi);
  const int sx_slot = 0, sy_slot = (D.axes & 1) ? 1 : 0, sz_slot = sy_slot + ((D.axes & 2) ? 1 : 0);

  auto load = [&](int kk) {
    VelPlane L{};
    const long tb = tb0 + (long)kk * sz, id = id0 + (long)kk * sz;
    if (x12.lds >= 0) L.hxx = ld(sxx, tb + x12.off);
    if (x21.lds >= 0) { L.hxy = ld(sxy, tb + x21.off); L.hxz = ld(sxz, tb + x21.off); }
    if (y12.lds >= 0) L.hyy = ld(syy, tb + y12.off);
    if (y21.lds >= 0) { L.hyxy = ld(sxy, tb + y21.off); L.hyz = ld(syz, tb + y21.off); }
    if (T.live) {
      L.sxx = ld(sxx, id); L.sxy = ld(sxy, id); L.syy = ld(syy, id);
      L.nzz = ld(szz, id + 2 * sz); L.nxz = ld(sxz, id + sz); L.nyz = ld(syz, id + sz);
      L.ovx = ld(fvx, id); L.ovy = ld(fvy, id); L.ovz = ld(fvz, id);
      L.dtr = fdtr[id];
      const long c = c00 + (long)plane * kk;
      const int ek = D.k0 + T.kc0 + kk;
      L.inz = (D.axes & 4) && (ek < D.zlo || ek >= D.zhi);
      if (inx) L.mx = psi_cell(psi + (long)(sx_slot * 6) * ncell + c, coef + (long)(sx_slot * 4) * ncell + c, ncell);
      if (iny) L.my = psi_cell(psi + (long)(sy_slot * 6) * ncell + c, coef + (long)(sy_slot * 4) * ncell + c, ncell);
      if (L.inz)
        L.mz = psi_cell(psi + (long)(sz_slot * 6) * ncell + c, coef + (long)(sz_slot * 4) * ncell + c, ncell);
    }
    return L;
  };
  // queues: qzz planes k-1 .. k+2 (szz forward-staggered), qxz, qyz planes k-2 .. k+1
  float qzz[4] = {0.f, 0.f, 0.f, 0.f}, qxz[4] = {0.f, 0.f, 0.f, 0.f}, qyz[4] = {0.f, 0.f, 0.f, 0.f};
  if (T.live) {  // one plane behind: fill() shifts plane 0's heads in
#pragma unroll
    for (int r = 1; r < 4; r++) {
      qzz[r] = ld(szz, id0 + (long)(r - 2) * sz);
      qxz[r] = ld(sxz, id0 + (long)(r - 3) * sz);
      qyz[r] = ld(syz, id0 + (long)(r - 3) * sz);
    }
  }
  VelPlane C = load(0);
  auto fill = [&](int buf, const VelPlane &L) {
    // queue heads of plane k enter the queues, then the tiles of plane k
#pragma unroll
    for (int r = 0; r < 3; r++) { qzz[r] = qzz[r + 1]; qxz[r] = qxz[r + 1]; qyz[r] = qyz[r + 1]; }
    qzz[3] = L.nzz; qxz[3] = L.nxz; qyz[3] = L.nyz;
    float *Lxx = lt[buf][0], *Lxy = lt[buf][1], *Lxz = lt[buf][2], *Lyy = lt[buf][3], *Lyz = lt[buf][4];
    if (T.live) {
      Lxx[T.me] = L.sxx; Lxy[T.me] = L.sxy; Lxz[T.me] = qxz[2]; Lyy[T.me] = L.syy; Lyz[T.me] = qyz[2];
    }
    if (x12.lds >= 0) Lxx[x12.lds] = L.hxx;
    if (x21.lds >= 0) { Lxy[x21.lds] = L.hxy; Lxz[x21.lds] = L.hxz; }
    if (y12.lds >= 0) Lyy[y12.lds] = L.hyy;
    if (y21.lds >= 0) { Lxy[y21.lds] = L.hyxy; Lyz[y21.lds] = L.hyz; }
  };
  fill(0, C);
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < P; kk++) {
    if (kk >= T.np) break;
    const int b = kk & 1;
    float nv0 = 0.f, nv1 = 0.f, nv2 = 0.f, np[9];
    if (T.live) {
      const float *Lxx = lt[b][0], *Lxy = lt[b][1], *Lxz = lt[b][2], *Lyy = lt[b][3], *Lyz = lt[b][4];
      const int me = T.me, px = T.px;
      float sxx_x = tp2(Lxx, me, 1, c0, c1), syx_x = tm2(Lxy, me, 1, c0, c1), szx_x = tm2(Lxz, me, 1, c0, c1);
      float syy_y = tp2(Lyy, me, px, c0, c1), sxy_y = tm2(Lxy, me, px, c0, c1), szy_y = tm2(Lyz, me, px, c0, c1);
      float szz_z = 0.f, sxz_z = 0.f, syz_z = 0.f;
      szz_z += c0 * (qzz[2] - qzz[1]);
      szz_z += c1 * (qzz[3] - qzz[0]);
      sxz_z += c0 * (qxz[2] - qxz[1]);
      sxz_z += c1 * (qxz[3] - qxz[0]);
      syz_z += c0 * (qyz[2] - qyz[1]);
      syz_z += c1 * (qyz[3] - qyz[0]);
      sxx_x = sxx_x / dx; syy_y = syy_y / dy; szz_z = szz_z / dz;
      sxy_y = sxy_y / dy; sxz_z = sxz_z / dz; syx_x = syx_x / dx;
      syz_z = syz_z / dz; szx_x = szx_x / dx; szy_y = szy_y / dy;
      if (inx) {
        const PsiCell<gfloat> &m = C.mx;
        np[0] = m.bh * m.m0 + m.ah * sxx_x;
        np[1] = m.b * m.m1 + m.a * syx_x;
        np[2] = m.b * m.m2 + m.a * szx_x;
        sxx_x += np[0]; syx_x += np[1]; szx_x += np[2];
      }
      if (iny) {
        const PsiCell<gfloat> &m = C.my;
        np[3] = m.b * m.m0 + m.a * sxy_y;
        np[4] = m.bh * m.m1 + m.ah * syy_y;
        np[5] = m.b * m.m2 + m.a * szy_y;
        sxy_y += np[3]; syy_y += np[4]; szy_y += np[5];
      }
      if (C.inz) {
        const PsiCell<gfloat> &m = C.mz;
        np[6] = m.b * m.m0 + m.a * sxz_z;
        np[7] = m.b * m.m1 + m.a * syz_z;
        np[8] = m.bh * m.m2 + m.ah * szz_z;
        sxz_z += np[6]; syz_z += np[7]; szz_z += np[8];
      }
      nv0 = C.ovx + C.dtr * (sxx_x + sxy_y + sxz_z);
      nv1 = C.ovy + C.dtr * (syx_x + syy_y + syz_z);
      nv2 = C.ovz + C.dtr * (szx_x + szy_y + szz_z);
    }
    const bool more = kk + 1 < T.np;
    VelPlane N{};
    if (more) N = load(kk + 1);
    if (T.live) {
      const long id = id0 + (long)kk * sz;
      const long c = c00 + (long)plane * kk;
      if (inx) {
        gfloat *q = psi + (long)(sx_slot * 6) * ncell + c;
        stq(q, 0, np[0]); stq(q, ncell, np[1]); stq(q, 2 * ncell, np[2]);
      }
      if (iny) {
        gfloat *q = psi + (long)(sy_slot * 6) * ncell + c;
        stq(q, 0, np[3]); stq(q, ncell, np[4]); stq(q, 2 * ncell, np[5]);
      }
      if (C.inz) {
        gfloat *q = psi + (long)(sz_slot * 6) * ncell + c;
        stq(q, 0, np[6]); stq(q, ncell, np[7]); stq(q, 2 * ncell, np[8]);
      }
      stq(fvx, id, nv0);
      stq(fvy, id, nv1);
      stq(fvz, id, nv2);
      if (D.mir) {
        const MirCell m = mir_cell(D.mir, ei, ej, D.k0 + T.kc0 + kk);
        if (m.any) { mir_put<gfloat>(m, VX, nv0); mir_put<gfloat>(m, VY, nv1); mir_put<gfloat>(m, VZ, nv2); }
      }
    }
    if (more) {
      fill(b ^ 1, N);
      __syncthreads();
      C = N;
    }
  }
}

// One plane's loads for a cell of the stress update.
struct StressPlane {
  float hxx, hxy, hyx, hyy, hzx, hzy;  // halo cells: vx (x, 2|1), (y, 1|2); vy (x, 1|2), (y, 2|1); vz (x, 1|2), (y, 1|2)
  float nx, ny, nz;                    // queue heads: vx and vy plane k + 2, vz plane k + 1
  float lm, mu, oxx, oyy, ozz, oxy, oxz, oyz, oau;
  PsiCell<gfloat> mx, my, mz;
  bool inz;
};

template <int WITH_SAU, int P>
__global__ __launch_bounds__(kBlock) DDX_PML_ATTR void k_stress_pml_t(const SlabDesc *__restrict__ descs,
                                                         const int *__restrict__ starts, int n, float dx,
                                                         float dy, float dz, float dt, int grp, PmlOrder ord) {
  __shared__ float lt[2][3][kPmlLds];  // vx, vy, vz
  int di, zc, tile;
  if (!pml_block(descs, starts, n, grp, ord, kPmlTiled + P, di, zc, tile)) return;
  const SlabDesc &D = descs[di];
  const PmlTile T = pml_tile(D, zc, tile, P);
  const long ncell = (long)D.ni * D.nj * D.nk;
  const int plane = D.ni * D.nj;
  const int sy = (int)D.sy;
  const long sz = D.sz;
  const long tb0 = (long)(D.i0 + T.x0) + D.sy * (D.j0 + T.y0) + sz * (D.k0 + T.kc0);
  const long id0 = tb0 + T.tx + (long)T.ty * sy;
  const long c00 = (long)(T.x0 + T.tx) + (long)D.ni * (T.y0 + T.ty) + (long)plane * T.kc0;
  const int ei = D.i0 + T.x0 + T.tx, ej = D.j0 + T.y0 + T.ty;
  const int t = (int)threadIdx.x;
  const TileSlot x12 = tile_slot(t, true, 1, 2, T.w, T.h, T.px, sy);
  const TileSlot x21 = tile_slot(t, true, 2, 1, T.w, T.h, T.px, sy);
  const TileSlot y12 = tile_slot(t, false, 1, 2, T.w, T.h, T.px, sy);
  const TileSlot y21 = tile_slot(t, false, 2, 1, T.w, T.h, T.px, sy);
  const float c0 = c_fd[2][0], c1 = c_fd[2][1];
  const gfloat *vx = (const gfloat *)D.f[VX], *vy = (const gfloat *)D.f[VY], *vz = (const gfloat *)D.f[VZ];
  gfloat *fsxx = (gfloat *)D.f[SXX], *fsyy = (gfloat *)D.f[SYY], *fszz = (gfloat *)D.f[SZZ];
  gfloat *fsxy = (gfloat *)D.f[SXY], *fsxz = (gfloat *)D.f[SXZ], *fsyz = (gfloat *)D.f[SYZ], *fsau = (gfloat *)D.f[SAU];
  gfloat *psi = (gfloat *)D.psi;
  const gfloat *flam = G(D.lam), *fmu = G(D.mu), *coef = G(D.coef);
  const bool inx = (D.axes & 1) && (ei < D.xlo || ei >= D.xhi);
  const bool iny = (D.axes & 2) && (ej < D.ylo || ej >= D.yhi);
  const int sx_slot = 0, sy_slot = (D.axes & 1) ? 1 : 0, sz_slot = sy_slot + ((D.axes & 2) ? 1 : 0);

  auto load = [&](int kk) {
    StressPlane L{};
    const long tb = tb0 + (long)kk * sz, id = id0 + (long)kk * sz;
    if (x21.lds >= 0) L.hxx = ld(vx, tb + x21.off);
    if (y12.lds >= 0) { L.hxy = ld(vx, tb + y12.off); L.hzy = ld(vz, tb + y12.off); }
    if (x12.lds >= 0) { L.hyx = ld(vy, tb + x12.off); L.hzx = ld(vz, tb + x12.off); }
    if (y21.lds >= 0) L.hyy = ld(vy, tb + y21.off);
    if (T.live) {
      L.nx = ld(vx, id + 2 * sz); L.ny = ld(vy, id + 2 * sz); L.nz = ld(vz, id + sz);
      L.lm = flam[id]; L.mu = fmu[id];
      L.oxx = ld(fsxx, id); L.oyy = ld(fsyy, id); L.ozz = ld(fszz, id);
      L.oxy = ld(fsxy, id); L.oxz = ld(fsxz, id); L.oyz = ld(fsyz, id);
      L.oau = WITH_SAU ? ld(fsau, id) : 0.f;
      const long c = c00 + (long)plane * kk;
      const int ek = D.k0 + T.kc0 + kk;
      L.inz = (D.axes & 4) && (ek < D.zlo || ek >= D.zhi);
      if (inx)
        L.mx = psi_cell(psi + (long)(sx_slot * 6 + 3) * ncell + c, coef + (long)(sx_slot * 4) * ncell + c, ncell);
      if (iny)
        L.my = psi_cell(psi + (long)(sy_slot * 6 + 3) * ncell + c, coef + (long)(sy_slot * 4) * ncell + c, ncell);
      if (L.inz)
        L.mz = psi_cell(psi + (long)(sz_slot * 6 + 3) * ncell + c, coef + (long)(sz_slot * 4) * ncell + c, ncell);
    }
    return L;
  };
  // queues: qx, qy planes k-1 .. k+2 (vx, vy forward-staggered in z), qz planes k-2 .. k+1
  float qx[4] = {0.f, 0.f, 0.f, 0.f}, qy[4] = {0.f, 0.f, 0.f, 0.f}, qz[4] = {0.f, 0.f, 0.f, 0.f};
  if (T.live) {  // one plane behind: fill() shifts plane 0's heads in
#pragma unroll
    for (int r = 1; r < 4; r++) {
      qx[r] = ld(vx, id0 + (long)(r - 2) * sz);
      qy[r] = ld(vy, id0 + (long)(r - 2) * sz);
      qz[r] = ld(vz, id0 + (long)(r - 3) * sz);
    }
  }
  StressPlane C = load(0);
  auto fill = [&](int buf, const StressPlane &L) {
#pragma unroll
    for (int r = 0; r < 3; r++) { qx[r] = qx[r + 1]; qy[r] = qy[r + 1]; qz[r] = qz[r + 1]; }
    qx[3] = L.nx; qy[3] = L.ny; qz[3] = L.nz;
    float *Lx = lt[buf][0], *Ly = lt[buf][1], *Lz = lt[buf][2];
    if (T.live) { Lx[T.me] = qx[1]; Ly[T.me] = qy[1]; Lz[T.me] = qz[2]; }
    if (x21.lds >= 0) Lx[x21.lds] = L.hxx;
    if (y12.lds >= 0) { Lx[y12.lds] = L.hxy; Lz[y12.lds] = L.hzy; }
    if (x12.lds >= 0) { Ly[x12.lds] = L.hyx; Lz[x12.lds] = L.hzx; }
    if (y21.lds >= 0) Ly[y21.lds] = L.hyy;
  };
  fill(0, C);
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < P; kk++) {
    if (kk >= T.np) break;
    const int b = kk & 1;
    float ns[7], np[9];
    if (T.live) {
      const float *Lx = lt[b][0], *Ly = lt[b][1], *Lz = lt[b][2];
      const int me = T.me, px = T.px;
      float vx_x = tm2(Lx, me, 1, c0, c1), vy_x = tp2(Ly, me, 1, c0, c1), vz_x = tp2(Lz, me, 1, c0, c1);
      float vy_y = tm2(Ly, me, px, c0, c1), vx_y = tp2(Lx, me, px, c0, c1), vz_y = tp2(Lz, me, px, c0, c1);
      float vz_z = 0.f, vx_z = 0.f, vy_z = 0.f;
      vz_z += c0 * (qz[2] - qz[1]);
      vz_z += c1 * (qz[3] - qz[0]);
      vx_z += c0 * (qx[2] - qx[1]);
      vx_z += c1 * (qx[3] - qx[0]);
      vy_z += c0 * (qy[2] - qy[1]);
      vy_z += c1 * (qy[3] - qy[0]);
      const float lm = C.lm, mu = C.mu;
      const float l2m = lm + 2.0f * mu;
      vx_x = vx_x / dx; vy_y = vy_y / dy; vz_z = vz_z / dz;
      vy_x = vy_x / dx; vz_x = vz_x / dx; vx_y = vx_y / dy;
      vz_y = vz_y / dy; vx_z = vx_z / dz; vy_z = vy_z / dz;
      if (inx) {
        const PsiCell<gfloat> &m = C.mx;
        np[0] = m.b * m.m0 + m.a * vx_x;
        np[1] = m.bh * m.m1 + m.ah * vy_x;
        np[2] = m.bh * m.m2 + m.ah * vz_x;
        vx_x += np[0]; vy_x += np[1]; vz_x += np[2];
      }
      if (iny) {
        const PsiCell<gfloat> &m = C.my;
        np[3] = m.bh * m.m0 + m.ah * vx_y;
        np[4] = m.b * m.m1 + m.a * vy_y;
        np[5] = m.bh * m.m2 + m.ah * vz_y;
        vx_y += np[3]; vy_y += np[4]; vz_y += np[5];
      }
      if (C.inz) {
        const PsiCell<gfloat> &m = C.mz;
        np[6] = m.bh * m.m0 + m.ah * vx_z;
        np[7] = m.bh * m.m1 + m.ah * vy_z;
        np[8] = m.b * m.m2 + m.a * vz_z;
        vx_z += np[6]; vy_z += np[7]; vz_z += np[8];
      }
      ns[0] = C.oxx + dt * (l2m * vx_x + lm * vy_y + lm * vz_z);
      ns[1] = C.oyy + dt * (lm * vx_x + l2m * vy_y + lm * vz_z);
      ns[2] = C.ozz + dt * (lm * vx_x + lm * vy_y + l2m * vz_z);
      ns[3] = WITH_SAU ? C.oau + dt * (l2m * vx_x + l2m * vy_y + l2m * vz_z) : 0.f;
      ns[4] = C.oxy + dt * mu * (vy_x + vx_y);
      ns[5] = C.oxz + dt * mu * (vz_x + vx_z);
      ns[6] = C.oyz + dt * mu * (vz_y + vy_z);
    }
    const bool more = kk + 1 < T.np;
    StressPlane N{};
    if (more) N = load(kk + 1);
    if (T.live) {
      const long id = id0 + (long)kk * sz;
      const long c = c00 + (long)plane * kk;
      if (inx) {
        gfloat *q = psi + (long)(sx_slot * 6 + 3) * ncell + c;
        stq(q, 0, np[0]); stq(q, ncell, np[1]); stq(q, 2 * ncell, np[2]);
      }
      if (iny) {
        gfloat *q = psi + (long)(sy_slot * 6 + 3) * ncell + c;
        stq(q, 0, np[3]); stq(q, ncell, np[4]); stq(q, 2 * ncell, np[5]);
      }
      if (C.inz) {
        gfloat *q = psi + (long)(sz_slot * 6 + 3) * ncell + c;
        stq(q, 0, np[6]); stq(q, ncell, np[7]); stq(q, 2 * ncell, np[8]);
      }
      stq(fsxx, id, ns[0]);
      stq(fsyy, id, ns[1]);
      stq(fszz, id, ns[2]);
      if (WITH_SAU) stq(fsau, id, ns[3]);
      stq(fsxy, id, ns[4]);
      stq(fsxz, id, ns[5]);
      stq(fsyz, id, ns[6]);
      if (D.mir) {
        const MirCell m = mir_cell(D.mir, ei, ej, D.k0 + T.kc0 + kk);
        if (m.any) {
          mir_put<gfloat>(m, SXX, ns[0]); mir_put<gfloat>(m, SYY, ns[1]); mir_put<gfloat>(m, SZZ, ns[2]);
          if (WITH_SAU) mir_put<gfloat>(m, SAU, ns[3]);
          mir_put<gfloat>(m, SXY, ns[4]); mir_put<gfloat>(m, SXZ, ns[5]); mir_put<gfloat>(m, SYZ, ns[6]);
        }
      }
    }
    if (more) {
      fill(b ^ 1, N);
      __syncthreads();
      C = N;
    }
  }
}

// ---------------------------------------------------------------------------
// CMM halo exchange between subdomains on one device
// (OP/jarvis_2_mpicu_halo_comm_operator_meat.hpp:299-326 extract into the
// send block, then the receiver's ghost layers): one copy per element from
// the sender's boundary layers to the receiver's ghosts, all links of a
// batch in one launch.  Boxes hold fewer than 2^31 elements (32-bit math).
// ---------------------------------------------------------------------------
// E: element type (float, or the 16-bit code of the 16-bit formats).  The
// copy walks rows (j, k) of the box, row index j + nj k.  Narrow boxes (the
// x-face links: 3-4 cells per row) take one cell per thread, a row's cells
// on consecutive lanes (measured: one row per thread, 45 us per exchange
// against 35, since each of its loads then touches 64 lines); wide boxes
// take one row per wavefront, lanes 64 cells apart, up to four loads in
// flight per lane before their stores, with no per-cell index division.
template <typename E>
__global__ __launch_bounds__(kBlock) void k_halo(const HaloDesc *__restrict__ descs,
                                                 const int *__restrict__ starts, int n) {
  typedef __attribute__((address_space(1))) E gE;
  const int b = blockIdx.x;
  const HaloDesc &D = descs[find_desc(starts, n, b)];
  const int rows = D.nj * D.nk;
  const gE *__restrict__ src = (const gE *)D.src;
  gE *__restrict__ dst = (gE *)D.dst;
  if (D.ni <= kHaloNarrow) {
    // consecutive lanes take consecutive cells, so the lanes of one row
    // share each load's cache line
    const int e = (b - D.blk0) * kBlock + (int)threadIdx.x;
    if (e >= rows * D.ni) return;
    const int row = e / D.ni, i = e - row * D.ni;
    const int k = row / D.nj, j = row - k * D.nj;
    const long s = (long)(D.si + i) + D.ssy * (D.sj + j) + D.ssz * (D.sk + k);
    const long d = (long)(D.di + i) + D.dsy * (D.dj + j) + D.dsz * (D.dk + k);
    dst[d] = src[s];
    return;
  }
  const int row = (b - D.blk0) * (kBlock / 64) + (int)(threadIdx.x >> 6);
  if (row >= rows) return;
  const int lane = (int)(threadIdx.x & 63);
  const int k = row / D.nj, j = row - k * D.nj;
  const long s = (long)D.si + D.ssy * (D.sj + j) + D.ssz * (D.sk + k);
  const long d = (long)D.di + D.dsy * (D.dj + j) + D.dsz * (D.dk + k);
  for (int i0 = 0; i0 < D.ni; i0 += 256) {
    E v[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int i = i0 + 64 * q + lane;
      if (i < D.ni) v[q] = src[s + i];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int i = i0 + 64 * q + lane;
      if (i < D.ni) dst[d + i] = v[q];
    }
  }
}

// ---------------------------------------------------------------------------
// Box <-> contiguous buffer copies (FLD/jarvis_3_device_field_meat.hpp:23-60):
// the VOWR boundary save / inject (BASE/fdtd3d_base_5_wave_recon_meat.hpp:89-133).  Boxes are walked
// x-fastest, so both sides of every copy are coalesced along x.
// ---------------------------------------------------------------------------
// Buffer element of box element (i, j, k): the box may be a sub-box of the
// buffer's box (CopyDesc::bi..bnj); for a whole box this is buf_off + the
// element's x-fastest index.
__device__ __forceinline__ long buf_index(const CopyDesc &D, int i, int j, int k) {
  return D.buf_off + (long)(D.bi + i) + (long)D.bni * ((long)(D.bj + j) + (long)D.bnj * (D.bk + k));
}

template <int DIR, typename E>
__global__ __launch_bounds__(kBlock) void k_copy(const CopyDesc *__restrict__ descs,
                                                 const int *__restrict__ starts, int n, void *__restrict__ buf) {
  typedef __attribute__((address_space(1))) E gE;
  const int b = blockIdx.x;
  const CopyDesc &D = descs[find_desc(starts, n, b)];
  // boxes hold fewer than 2^31 elements (checked on the host): 32-bit index math
  const int cnt = D.ni * D.nj * D.nk;
  const int e = (b - D.blk0) * kBlock + (int)threadIdx.x;
  if (e >= cnt) return;
  const int r = e / D.ni;
  const int k = r / D.nj;
  const int i = e - r * D.ni;
  const int j = r - k * D.nj;
  const long id = (long)(D.i0 + i) + D.sy * (D.j0 + j) + D.sz * (D.k0 + k);
  const long bo = buf_index(D, i, j, k);
  if (DIR > 0) ((gE *)buf)[bo] = ((const gE *)D.field)[id];
  else ((gE *)D.field)[id] = ((const gE *)buf)[bo];
}

// The same boxes between FP32 fields and a 16-bit VOWR store (S: F16 or
// BF16): saved values are the field times the class scale, rounded to
// nearest even; injected values are widened and divided by it (exact:
// powers of two).
template <int DIR, class S>
__global__ __launch_bounds__(kBlock) void k_copy_conv(const CopyDesc *__restrict__ descs,
                                                      const int *__restrict__ starts, int n, void *__restrict__ buf,
                                                      float sc_v, float sc_s) {
  typedef typename GPtr<S>::T gT;
  const int b = blockIdx.x;
  const CopyDesc &D = descs[find_desc(starts, n, b)];
  const int cnt = D.ni * D.nj * D.nk;
  const int e = (b - D.blk0) * kBlock + (int)threadIdx.x;
  if (e >= cnt) return;
  const int r = e / D.ni;
  const int k = r / D.nj;
  const int i = e - r * D.ni;
  const int j = r - k * D.nj;
  const long id = (long)(D.i0 + i) + D.sy * (D.j0 + j) + D.sz * (D.k0 + k);
  const float sc = D.cls == 0 ? sc_v : sc_s;
  gT *ring = (gT *)buf;
  const long bo = buf_index(D, i, j, k);
  if (DIR > 0) stq(ring, bo, G(D.field)[id] * sc);
  else G(D.field)[id] = ld(ring, bo) * (1.0f / sc);
}

// Per-slice scaling of the 16-bit VOWR store: the largest |value| of each
// ring field over its boxes (block maximum, then one atomic per block), the
// slice's exponents, and the copies with them.  Scaling by a power of two
// and its inverse is exact, so a slice's values are rounded once, to the
// 16-bit format, as in the fixed-scale copy.
__global__ __launch_bounds__(kBlock) void k_ring_max(const CopyDesc *__restrict__ descs,
                                                     const int *__restrict__ starts, int n, unsigned *maxbits) {
  __shared__ unsigned red[kBlock];
  const int b = blockIdx.x;
  const CopyDesc &D = descs[find_desc(starts, n, b)];
  const int cnt = D.ni * D.nj * D.nk;
  const int e = (b - D.blk0) * kBlock + (int)threadIdx.x;
  unsigned m = 0;
  if (e < cnt) {
    const int r = e / D.ni;
    const int k = r / D.nj;
    const int i = e - r * D.ni;
    const int j = r - k * D.nj;
    const long id = (long)(D.i0 + i) + D.sy * (D.j0 + j) + D.sz * (D.k0 + k);
    m = __float_as_uint(fabsf(G(D.field)[id]));  // non-negative floats order as their bits
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] != 0u) atomicMax(maxbits + D.fid, red[0]);
}

__global__ void k_ring_exp(unsigned *maxbits, int *exps, int nf) {
  const int f = threadIdx.x;
  if (f >= nf) return;
  const float m = __uint_as_float(maxbits[f]);
  int e = 0;
  if (m > 0.f) {
    int ex = 0;
    (void)frexpf(m, &ex);
    e = min(100, max(-100, 15 - ex));
  }
  exps[f] = e;
  maxbits[f] = 0u;
}

template <int DIR, class S>
__global__ __launch_bounds__(kBlock) void k_copy_dyn(const CopyDesc *__restrict__ descs,
                                                     const int *__restrict__ starts, int n, void *__restrict__ buf,
                                                     const int *__restrict__ exps) {
  typedef typename GPtr<S>::T gT;
  const int b = blockIdx.x;
  const CopyDesc &D = descs[find_desc(starts, n, b)];
  const int cnt = D.ni * D.nj * D.nk;
  const int e = (b - D.blk0) * kBlock + (int)threadIdx.x;
  if (e >= cnt) return;
  const int r = e / D.ni;
  const int k = r / D.nj;
  const int i = e - r * D.ni;
  const int j = r - k * D.nj;
  const long id = (long)(D.i0 + i) + D.sy * (D.j0 + j) + D.sz * (D.k0 + k);
  const float sc = ldexpf(1.0f, exps[D.fid]);
  gT *ring = (gT *)buf;
  const long bo = buf_index(D, i, j, k);
  if (DIR > 0) stq(ring, bo, G(D.field)[id] * sc);
  else G(D.field)[id] = ld(ring, bo) * (1.0f / sc);
}

// Block-scaled 16-bit VOWR store over FP32 wavefields (STORE_BFP16).  The
// entries of a slice form runs of kRingBlock, each ring box starting on a
// run; a run shares the exponent s of its largest |value| (largest =
// m 2^s, 0.5 <= m < 1, one byte; s = -126 for a zero or subnormal run).  An
// entry codes u = |v| 2^-s, in [0, 1), as a sign and a 15-bit magnitude c:
// for u >= 2^-7, c = 4096 E + M with u = (4096 + M) 2^(E - 20), E = 1..7 the
// binade and M a 12-bit mantissa (rounding error at most 2^-13 of u, FP16's
// is 2^-11); below, c = u 2^19.  The codes are monotonic in u, so a mantissa
// that rounds up carries into the exponent.  Every scaling is by a power of
// two and exact.
typedef __attribute__((address_space(1))) short gshort;

__device__ __forceinline__ short bfp_code(float v, int s) {
  const float u = ldexpf(fabsf(v), -s);
  int c;
  if (u >= 0.0078125f) {  // 2^-7
    int ex = 0;
    (void)frexpf(u, &ex);  // u in [2^(ex-1), 2^ex), ex in -6..0
    c = 4096 * (ex + 6) + (int)rintf(ldexpf(u, 13 - ex));
  } else {
    c = (int)rintf(ldexpf(u, 19));
  }
  c = min(c, 32767);
  return (short)(v < 0.f ? -c : c);
}

__device__ __forceinline__ float bfp_value(short q, int s) {
  const int c = q < 0 ? -(int)q : (int)q;
  const float u = c < 4096 ? ldexpf((float)c, -19) : ldexpf((float)(4096 + (c & 4095)), (c >> 12) - 20);
  return ldexpf(q < 0 ? -u : u, s);
}

// Save: the save table holds whole boxes (buffer index = buf_off + the
// cell's x-fastest index e, buf_off a multiple of kRingBlock) and a block
// covers kBlock consecutive e of one box, so a run is exactly kRingBlock
// consecutive lanes: their largest |value| comes from lane shuffles (every
// lane takes part; cells past the box count as 0), and the run's first lane
// writes its exponent.
static_assert(kBlock % kRingBlock == 0 && 64 % kRingBlock == 0, "runs must tile the wavefronts");
__global__ __launch_bounds__(kBlock) void k_save_bfp(const CopyDesc *__restrict__ descs,
                                                     const int *__restrict__ starts, int n, void *__restrict__ buf,
                                                     signed char *__restrict__ bexp) {
  const int b = blockIdx.x;
  const CopyDesc &D = descs[find_desc(starts, n, b)];
  const int cnt = D.ni * D.nj * D.nk;
  const int e = (b - D.blk0) * kBlock + (int)threadIdx.x;
  const bool in = e < cnt;
  float v = 0.f;
  long id = 0;
  if (in) {
    const int r = e / D.ni;
    const int k = r / D.nj;
    const int i = e - r * D.ni;
    const int j = r - k * D.nj;
    id = (long)(D.i0 + i) + D.sy * (D.j0 + j) + D.sz * (D.k0 + k);
    v = G(D.field)[id];
  }
  unsigned m = __float_as_uint(fabsf(v));  // non-negative floats order as their bits
#pragma unroll
  for (int w = kRingBlock / 2; w > 0; w >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, w, kRingBlock));
  const float mx = __uint_as_float(m);
  int s = -126;
  if (mx >= 1.17549435e-38f && mx <= 3.4028235e38f) {
    (void)frexpf(mx, &s);
    s = max(s, -126);
  }
  if (!in) return;
  const long bo = D.buf_off + e;
  ((gshort *)buf)[bo] = bfp_code(v, s);
  if ((threadIdx.x & (kRingBlock - 1)) == 0) bexp[bo / kRingBlock] = (signed char)s;
}

// Inject (any ring table, sub-boxes included).
__global__ __launch_bounds__(kBlock) void k_inject_bfp(const CopyDesc *__restrict__ descs,
                                                       const int *__restrict__ starts, int n,
                                                       const void *__restrict__ buf,
                                                       const signed char *__restrict__ bexp) {
  const int b = blockIdx.x;
  const CopyDesc &D = descs[find_desc(starts, n, b)];
  const int cnt = D.ni * D.nj * D.nk;
  const int e = (b - D.blk0) * kBlock + (int)threadIdx.x;
  if (e >= cnt) return;
  const int r = e / D.ni;
  const int k = r / D.nj;
  const int i = e - r * D.ni;
  const int j = r - k * D.nj;
  const long id = (long)(D.i0 + i) + D.sy * (D.j0 + j) + D.sz * (D.k0 + k);
  const long bo = buf_index(D, i, j, k);
  G(D.field)[id] = bfp_value(((const gshort *)buf)[bo], bexp[bo / kRingBlock]);
}

// Explosive source (PHY/fdtd3d_phy_mpi_3_stress_kernel.hpp:1308-1364); one
// thread, launched only for the subdomain that owns the shot node.
// ss: stress-class scale of the 16-bit formats (1 for F32).  The time index
// is it, or *itp when itp is set (a device step counter: launches captured
// once in a graph and replayed every step).
template <class S>
__global__ void k_source(void *psxx, void *psyy, void *pszz, long off, const float *w, int it, float sign,
                         int src_mode, float ss, const int *itp) {
  typedef typename GPtr<S>::T gT;
  gT *sxx = (gT *)psxx, *syy = (gT *)psyy, *szz = (gT *)pszz;
  if (itp) it = *itp;
  const float v = std::is_same<S, F32>::value ? w[it] : w[it] * ss;
  if (sign > 0) {
    if (src_mode == 0) { add_to(sxx, off, v); add_to(syy, off, v); }
    add_to(szz, off, v);
  } else {
    if (src_mode == 0) { add_to(sxx, off, -v); add_to(syy, off, -v); }
    add_to(szz, off, -v);
  }
}

// The source of every member of a fused batch (one thread per member), at
// the member's own time index; the same operations as k_source.
template <class S>
__global__ void k_source_batch(const MemberIO *__restrict__ io, int n, int gs, const int *__restrict__ gsp,
                               float sign) {
  typedef typename GPtr<S>::T gT;
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const MemberIO &M = io[q];
  const int g = gsp ? *gsp : gs;
  const int it = M.rev ? M.join - g : g - M.join;
  if (!M.sxx || it < 0 || it >= M.ntime) return;
  // sign = -1 subtracts (the RTM backward pass): x + (-w) is x - w exactly
  const float v = (std::is_same<S, F32>::value ? M.wavelet[it] : M.wavelet[it] * M.ss) * sign;
  if (M.src_mode == 0) {
    add_to((gT *)M.sxx, M.src_off, v);
    add_to((gT *)M.syy, M.src_off, v);
  }
  add_to((gT *)M.szz, M.src_off, v);
}

// Receiver samples of every member of a fused batch: blocks of kBlock
// receivers, each block within one member; the same operations as
// k_recv_sample.
template <class S>
__global__ __launch_bounds__(kBlock) void k_recv_batch(const MemberIO *__restrict__ io, const int *__restrict__ starts,
                                                       int n, int gs, const int *__restrict__ gsp) {
  typedef typename GPtr<S>::T gT;
  const int b = blockIdx.x;
  const MemberIO &M = io[find_desc(starts, n, b)];
  const int r = (b - M.blk0) * kBlock + (int)threadIdx.x;
  const int it = (gsp ? *gsp : gs) - M.join;
  if (r >= M.nrec || it < 0 || it >= M.ntime) return;
  const int s = M.rsub[r];
  const long o = M.roff[r];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const float v = ld((const gT *)M.vptr[s * 3 + q], o);
    M.seis[((long)q * M.nrec + r) * M.ntime + it] = std::is_same<S, F32>::value ? v : v * M.inv_sv;
  }
}

// Reverse-time receiver injection of every member of a fused RTM batch
// (the observed data M.seis at the member's time index into the velocities
// M.vptr): the same operations as k_recv_inject.
template <class S>
__global__ __launch_bounds__(kBlock) void k_inject_batch(const MemberIO *__restrict__ io,
                                                         const int *__restrict__ starts, int n, int gs) {
  typedef typename GPtr<S>::T gT;
  const int b = blockIdx.x;
  const MemberIO &M = io[find_desc(starts, n, b)];
  const int r = (b - M.blk0) * kBlock + (int)threadIdx.x;
  const int it = M.rev ? M.join - gs : gs - M.join;
  if (r >= M.nrec || it < 0 || it >= M.ntime) return;
  const int s = M.rsub[r];
  const long o = M.roff[r];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const float v = M.seis[((long)q * M.nrec + r) * M.ntime + it];
    add_to((gT *)M.vptr[s * 3 + q], o, std::is_same<S, F32>::value ? v : v * M.sv);
  }
}

// Receiver sampling (BASE/fdtd3d_base_4_seisrecord_meat.hpp:57-77) and
// reverse-time injection (:79-107), grid-stride over all kept receivers of
// all subdomains (no 1024-receiver limit).  seis: [3][nrec][ntime].
// sv: velocity-class scale of the 16-bit formats (1 for F32); samples are
// returned unscaled (exact: the scale is a power of two).
template <class S>
__global__ void k_recv_sample(void *const *__restrict__ vptr, const int *__restrict__ rsub,
                              const long *__restrict__ roff, int nrec, int it, int ntime,
                              float *__restrict__ seis, float inv_sv, const int *itp) {
  typedef typename GPtr<S>::T gT;
  if (itp) it = *itp;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrec; r += gridDim.x * blockDim.x) {
    const int s = rsub[r];
    const long o = roff[r];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const float v = ld((const gT *)vptr[s * 3 + q], o);
      seis[((long)q * nrec + r) * ntime + it] = std::is_same<S, F32>::value ? v : v * inv_sv;
    }
  }
}
template <class S>
__global__ void k_recv_inject(void *const *__restrict__ vptr, const int *__restrict__ rsub,
                              const long *__restrict__ roff, int nrec, int it, int ntime,
                              const float *__restrict__ seis, float sv) {
  typedef typename GPtr<S>::T gT;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrec; r += gridDim.x * blockDim.x) {
    const int s = rsub[r];
    const long o = roff[r];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const float v = seis[((long)q * nrec + r) * ntime + it];
      add_to((gT *)vptr[s * 3 + q], o, std::is_same<S, F32>::value ? v : v * sv);
    }
  }
}

// FWI adjoint source (BASE/fdtd3d_base_4_seisrecord_meat.hpp:109-147,
// simulate_type fwi): v += obs - syn at each receiver node.
__global__ void k_recv_residual(void *const *__restrict__ vptr, const int *__restrict__ rsub,
                                const long *__restrict__ roff, int nrec, int it, int ntime,
                                const float *__restrict__ obs, const float *__restrict__ syn) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrec; r += gridDim.x * blockDim.x) {
    const int s = rsub[r];
    const long o = roff[r];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const long k = ((long)q * nrec + r) * ntime + it;
      G(vptr[s * 3 + q])[o] += obs[k] - syn[k];
    }
  }
}

// ---------------------------------------------------------------------------
// Imaging: P/S-decomposed zero-lag cross-correlation
// (IMG/rtm3d_iamge_kernel.hpp:6-51), accumulated in T (float as the
// reference, or double).
// ---------------------------------------------------------------------------
// gridphy cell of this thread; a subdomain's gridphy holds fewer than 2^31
// cells (checked on the host), so the index math is 32-bit.
__device__ __forceinline__ void img_cell(const ImgDesc &D, int b, long &q, long &id, bool &ok) {
  const int np = D.nx * D.ny * D.nz;
  const int qi = (b - D.blk0) * kBlock + (int)threadIdx.x;
  ok = qi < np;
  const int r = qi / D.nx;
  const int k = r / D.ny;
  const int i = qi - r * D.nx;
  const int j = r - k * D.ny;
  q = qi;
  id = (long)(D.ox + i) + D.sy * (D.oy + j) + D.sz * (D.oz + k);
}

// S: storage format of the velocity fields; inv_sv: inverse velocity-class
// scale (values are unscaled before the products, exactly).
template <class S>
__device__ __forceinline__ float ld_v(const void *p, long id, float inv_sv) {
  const float v = ld_stream((const typename GPtr<S>::T *)p, id);
  return std::is_same<S, F32>::value ? v : v * inv_sv;
}
template <typename T, class S>
__global__ __launch_bounds__(kBlock) void k_image(const ImgDesc *__restrict__ descs,
                                                  const int *__restrict__ starts, int n, float inv_sv) {
  const int b = blockIdx.x;
  const ImgDesc &D = descs[find_desc(starts, n, b)];
  long q, id;
  bool ok;
  img_cell(D, b, q, id, ok);
  if (!ok) return;
  const float pfx = ld_v<S>(D.b[3], id, inv_sv), pfy = ld_v<S>(D.b[4], id, inv_sv), pfz = ld_v<S>(D.b[5], id, inv_sv);
  const float sfx = ld_v<S>(D.b[0], id, inv_sv) - pfx, sfy = ld_v<S>(D.b[1], id, inv_sv) - pfy,
              sfz = ld_v<S>(D.b[2], id, inv_sv) - pfz;
  const float prx = ld_v<S>(D.r[3], id, inv_sv), pry = ld_v<S>(D.r[4], id, inv_sv), prz = ld_v<S>(D.r[5], id, inv_sv);
  const float srx = ld_v<S>(D.r[0], id, inv_sv) - prx, sry = ld_v<S>(D.r[1], id, inv_sv) - pry,
              srz = ld_v<S>(D.r[2], id, inv_sv) - prz;
  // (the images keep read-add-store per image: this kernel runs at high
  // occupancy, and reading the five images first measured slower)
  typedef __attribute__((address_space(1))) T gT;
  gT *Ip = (gT *)D.img[0], *Ipp = (gT *)D.img[1], *Ips = (gT *)D.img[2], *Isp = (gT *)D.img[3], *Iss = (gT *)D.img[4];
  if constexpr (sizeof(T) == sizeof(float)) {
    rmw_add_t(Ip, q, pfx * pfx + pfy * pfy + pfz * pfz);
    rmw_add_t(Ipp, q, pfx * prx + pfy * pry + pfz * prz);
    rmw_add_t(Ips, q, pfx * srx + pfy * sry + pfz * srz);
    rmw_add_t(Isp, q, sfx * prx + sfy * pry + sfz * prz);
    rmw_add_t(Iss, q, sfx * srx + sfy * sry + sfz * srz);
  } else {
    const T PX = pfx, PY = pfy, PZ = pfz, SX = sfx, SY = sfy, SZ = sfz;
    const T px = prx, py = pry, pz = prz, sx = srx, sy = sry, sz = srz;
    rmw_add_t(Ip, q, PX * PX + PY * PY + PZ * PZ);
    rmw_add_t(Ipp, q, PX * px + PY * py + PZ * pz);
    rmw_add_t(Ips, q, PX * sx + PY * sy + PZ * sz);
    rmw_add_t(Isp, q, SX * px + SY * py + SZ * pz);
    rmw_add_t(Iss, q, SX * sx + SY * sy + SZ * sz);
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
// Instantiated z-blocked brick shapes (tx, ty, cy, cz): all of them for FP32
// storage, the fastest few for the 16-bit formats.
#define DDX_TILINGS(X)                                                                        \
  X(64, 4, 1, 2) X(64, 4, 1, 4) X(64, 4, 1, 8) X(128, 2, 1, 4) X(64, 8, 1, 4) X(32, 8, 1, 4) X(64, 4, 2, 4) \
  X(128, 4, 1, 2) X(64, 2, 1, 8)
#define DDX_TILINGS16(X) X(64, 4, 1, 2) X(64, 4, 1, 4) X(64, 4, 1, 8) X(128, 2, 1, 4) X(64, 2, 1, 8)
// paired bricks (px = 2), every storage format
#define DDX_TILINGS2(X) X(64, 4, 1, 1) X(64, 4, 1, 2) X(64, 4, 1, 4) X(64, 2, 1, 4)
// bricks with the y arm from LDS (xl = 2: tx = 64, cy = 1), (ty, kz), FP32
#define DDX_TILINGSY(X) X(4, 2) X(4, 4) X(8, 2) X(8, 4) X(16, 2) X(16, 4)
// lane-exchange bricks (xl: x arm from neighbouring lanes, tx = 64), FP32
#define DDX_TILINGSX(X) X(64, 4, 1, 2) X(64, 4, 1, 4) X(64, 4, 1, 8) X(64, 8, 1, 4) X(64, 2, 1, 8)

bool tiling_supported(const Tiling &t, int fmt) {
  if (t.xl == 2) {
    if (t.px != 1 || fmt != STORE_F32 || t.tx != 64 || t.cy != 1) return false;
#define DDX_T(TY, CZ) if (t.ty == TY && t.kz == CZ) return true;
    DDX_TILINGSY(DDX_T)
#undef DDX_T
    return false;
  }
  if (t.xl) {
    if (t.px != 1 || fmt != STORE_F32) return false;
#define DDX_T(TX, TY, CY, CZ) if (t.tx == TX && t.ty == TY && t.cy == CY && t.kz == CZ) return true;
    DDX_TILINGSX(DDX_T)
#undef DDX_T
    return false;
  }
  if (t.px == 2) {
#define DDX_T(TX, TY, CY, CZ) if (t.tx == TX && t.ty == TY && t.cy == CY && t.kz == CZ) return true;
    DDX_TILINGS2(DDX_T)
#undef DDX_T
    return false;
  }
  if (t.px != 1) return false;
  if (fmt != STORE_F32) {
#define DDX_T(TX, TY, CY, CZ) if (t.tx == TX && t.ty == TY && t.cy == CY && t.kz == CZ) return true;
    DDX_TILINGS16(DDX_T)
#undef DDX_T
    return false;
  }
  if (t.tx == LX && t.ty == LY && t.cy == 0) return t.kz > 0;  // LDS-tiled 2.5D kernels
#define DDX_T(TX, TY, CY, CZ) if (t.tx == TX && t.ty == TY && t.cy == CY && t.kz == CZ) return true;
  DDX_TILINGS(DDX_T)
#undef DDX_T
  return false;
}

int tiling_list(int fmt, Tiling *out, int max) {
  int n = 0;
  auto add = [&](int tx, int ty, int cy, int kz, int px, int xl = 0) {
    if (n < max) {
      out[n].tx = tx; out[n].ty = ty; out[n].cy = cy; out[n].kz = kz; out[n].px = px; out[n].xl = xl;
    }
    n++;
  };
#define DDX_T(TX, TY, CY, CZ) add(TX, TY, CY, CZ, 1);
  if (fmt == STORE_F32) {
    DDX_TILINGS(DDX_T)
  } else {
    DDX_TILINGS16(DDX_T)
  }
#undef DDX_T
#define DDX_T(TX, TY, CY, CZ) add(TX, TY, CY, CZ, 2);
  DDX_TILINGS2(DDX_T)
#undef DDX_T
  if (fmt == STORE_F32) {
#define DDX_T(TX, TY, CY, CZ) add(TX, TY, CY, CZ, 1, 1);
    DDX_TILINGSX(DDX_T)
#undef DDX_T
#define DDX_T(TY, CZ) add(64, TY, 1, CZ, 1, 2);
    DDX_TILINGSY(DDX_T)
#undef DDX_T
  }
  return n;
}

template <int TX, int TY, int CY, int CZ, class S, bool XL = false>
static void vel_zb(int mode, dim3 g, hipStream_t st, const PhyDesc *d, const int *starts, int n, float dx, float dy,
                   float dz, float sign, int gnx, int gny, int gnz, int R) {
  dim3 bl(TX * TY);
  switch (mode) {
    case PURE_FORWARD:
    case RTM_FORWARD:
      hipLaunchKernelGGL((k_vel_phy_zb<PURE_FORWARD, TX, TY, CY, CZ, S, XL>), g, bl, 0, st, d, starts, n, dx, dy, dz, sign,
                         gnx, gny, gnz, R);
      break;
    case RTM_BACKWARD:
      hipLaunchKernelGGL((k_vel_phy_zb<RTM_BACKWARD, TX, TY, CY, CZ, S, XL>), g, bl, 0, st, d, starts, n, dx, dy, dz, sign,
                         gnx, gny, gnz, R);
      break;
    case FWI:
      if constexpr (std::is_same<S, F32>::value && !XL)
        hipLaunchKernelGGL((k_vel_phy_zb<FWI, TX, TY, CY, CZ, F32>), g, bl, 0, st, d, starts, n, dx, dy, dz, sign, gnx,
                           gny, gnz, R);
      break;
    default:
      hipLaunchKernelGGL((k_vel_phy_zb<RTM_REVERSE, TX, TY, CY, CZ, S, XL>), g, bl, 0, st, d, starts, n, dx, dy, dz, sign,
                         gnx, gny, gnz, R);
  }
}

template <int TX, int TY, int CY, int CZ, class S, bool XL = false>
static void stress_zb(int mode, dim3 g, hipStream_t st, const PhyDesc *d, const int *starts, int n, float dx,
                      float dy, float dz, float dt, int gnx, int gny, int gnz, int R) {
  dim3 bl(TX * TY);
  switch (mode) {
    case PURE_FORWARD:
      hipLaunchKernelGGL((k_stress_phy_zb<PURE_FORWARD, TX, TY, CY, CZ, S, XL>), g, bl, 0, st, d, starts, n, dx, dy, dz,
                         dt, gnx, gny, gnz, R);
      break;
    case RTM_FORWARD:
      hipLaunchKernelGGL((k_stress_phy_zb<RTM_FORWARD, TX, TY, CY, CZ, S, XL>), g, bl, 0, st, d, starts, n, dx, dy, dz,
                         dt, gnx, gny, gnz, R);
      break;
    case RTM_BACKWARD:
      hipLaunchKernelGGL((k_stress_phy_zb<RTM_BACKWARD, TX, TY, CY, CZ, S, XL>), g, bl, 0, st, d, starts, n, dx, dy, dz,
                         dt, gnx, gny, gnz, R);
      break;
    case FWI:
      if constexpr (std::is_same<S, F32>::value && !XL)
        hipLaunchKernelGGL((k_stress_phy_zb<FWI, TX, TY, CY, CZ, F32>), g, bl, 0, st, d, starts, n, dx, dy, dz, dt,
                           gnx, gny, gnz, R);
      break;
    default:
      hipLaunchKernelGGL((k_stress_phy_zb<RTM_REVERSE, TX, TY, CY, CZ, S, XL>), g, bl, 0, st, d, starts, n, dx, dy, dz,
                         dt, gnx, gny, gnz, R);
  }
}

template <int TX, int TY, int CY, int CZ, class S>
static void vel_zb2(int mode, dim3 g, hipStream_t st, const PhyDesc *d, const int *starts, int n, float dx,
                    float dy, float dz, float sign, int gnx, int gny, int gnz, int R) {
  dim3 bl(TX * TY);
  switch (mode) {
    case PURE_FORWARD:
    case RTM_FORWARD:
      hipLaunchKernelGGL((k_vel_phy_zb2<PURE_FORWARD, TX, TY, CY, CZ, S>), g, bl, 0, st, d, starts, n, dx, dy, dz,
                         sign, gnx, gny, gnz, R);
      break;
    case RTM_BACKWARD:
      hipLaunchKernelGGL((k_vel_phy_zb2<RTM_BACKWARD, TX, TY, CY, CZ, S>), g, bl, 0, st, d, starts, n, dx, dy, dz,
                         sign, gnx, gny, gnz, R);
      break;
    case RTM_REVERSE:
      hipLaunchKernelGGL((k_vel_phy_zb2<RTM_REVERSE, TX, TY, CY, CZ, S>), g, bl, 0, st, d, starts, n, dx, dy, dz,
                         sign, gnx, gny, gnz, R);
      break;
    default:
      break;  // FWI: one-cell bricks only
  }
}

template <int TX, int TY, int CY, int CZ, class S>
static void stress_zb2(int mode, dim3 g, hipStream_t st, const PhyDesc *d, const int *starts, int n, float dx,
                       float dy, float dz, float dt, int gnx, int gny, int gnz, int R) {
  dim3 bl(TX * TY);
  switch (mode) {
    case PURE_FORWARD:
      hipLaunchKernelGGL((k_stress_phy_zb2<PURE_FORWARD, TX, TY, CY, CZ, S>), g, bl, 0, st, d, starts, n, dx, dy, dz,
                         dt, gnx, gny, gnz, R);
      break;
    case RTM_FORWARD:
      hipLaunchKernelGGL((k_stress_phy_zb2<RTM_FORWARD, TX, TY, CY, CZ, S>), g, bl, 0, st, d, starts, n, dx, dy, dz,
                         dt, gnx, gny, gnz, R);
      break;
    case RTM_BACKWARD:
      hipLaunchKernelGGL((k_stress_phy_zb2<RTM_BACKWARD, TX, TY, CY, CZ, S>), g, bl, 0, st, d, starts, n, dx, dy,
                         dz, dt, gnx, gny, gnz, R);
      break;
    case RTM_REVERSE:
      hipLaunchKernelGGL((k_stress_phy_zb2<RTM_REVERSE, TX, TY, CY, CZ, S>), g, bl, 0, st, d, starts, n, dx, dy, dz,
                         dt, gnx, gny, gnz, R);
      break;
    default:
      break;  // FWI: one-cell bricks only
  }
}

// kz == 0 selects the one-plane-per-block kernels (k_vel_phy / k_stress_phy),
// cy == 0 the LDS-tiled ones (both FP32 storage only).  FWI mode and the
// 16-bit formats run on the z-blocked bricks only (the engine never selects
// another shape for them).
bool launch_vel_phy(const Tiling &t, int mode, int fmt, const PhyDesc *d, const int *starts, int n, int nblk,
                    float dx, float dy, float dz, float sign, int gnx, int gny, int gnz, int R, hipStream_t st) {
  if (nblk <= 0) return true;
  if ((mode == FWI || fmt != STORE_F32) && !(t.kz > 0 && t.cy > 0)) return false;
  if (mode == FWI && (t.px != 1 || t.xl)) return false;
  dim3 g(nblk), bl(kBlock);
  if (t.xl == 2) {
    if (fmt != STORE_F32 || t.tx != 64 || t.cy != 1) return false;
#define DDX_T(TY, CZ)                                                                                  \
  if (t.ty == TY && t.kz == CZ) {                                                                      \
    const dim3 b2(64 * TY);                                                                            \
    switch (mode) {                                                                                    \
      case PURE_FORWARD:                                                                               \
      case RTM_FORWARD:                                                                                \
        hipLaunchKernelGGL((k_vel_phy_yl<PURE_FORWARD, TY, CZ, F32>), g, b2, 0, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R); \
        break;                                                                                         \
      case RTM_BACKWARD:                                                                               \
        hipLaunchKernelGGL((k_vel_phy_yl<RTM_BACKWARD, TY, CZ, F32>), g, b2, 0, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R); \
        break;                                                                                         \
      default:                                                                                         \
        hipLaunchKernelGGL((k_vel_phy_yl<RTM_REVERSE, TY, CZ, F32>), g, b2, 0, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R); \
    }                                                                                                  \
    return true;                                                                                      \
  }
    DDX_TILINGSY(DDX_T)
#undef DDX_T
    return false;
  }
  if (t.xl) {
    if (fmt != STORE_F32) return false;
#define DDX_T(TX, TY, CY, CZ)                                                                         \
  if (t.tx == TX && t.ty == TY && t.cy == CY && t.kz == CZ) {                                         \
    vel_zb<TX, TY, CY, CZ, F32, true>(mode, g, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R);  \
    return true;                                                                                     \
  }
    DDX_TILINGSX(DDX_T)
#undef DDX_T
    return false;
  }
  if (t.px == 2) {
#define DDX_T(TX, TY, CY, CZ)                                                                          \
  if (t.tx == TX && t.ty == TY && t.cy == CY && t.kz == CZ) {                                          \
    if (fmt == STORE_F16)                                                                              \
      vel_zb2<TX, TY, CY, CZ, F16>(mode, g, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R);     \
    else if (fmt == STORE_BF16)                                                                        \
      vel_zb2<TX, TY, CY, CZ, BF16>(mode, g, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R);    \
    else                                                                                               \
      vel_zb2<TX, TY, CY, CZ, F32>(mode, g, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R);     \
    return true;                                                                                      \
  }
    DDX_TILINGS2(DDX_T)
#undef DDX_T
    return false;
  }
  if (fmt == STORE_F16 || fmt == STORE_BF16) {
#define DDX_T(TX, TY, CY, CZ)                                                                         \
  if (t.tx == TX && t.ty == TY && t.cy == CY && t.kz == CZ) {                                         \
    if (fmt == STORE_F16)                                                                             \
      vel_zb<TX, TY, CY, CZ, F16>(mode, g, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R);     \
    else                                                                                              \
      vel_zb<TX, TY, CY, CZ, BF16>(mode, g, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R);    \
    return true;                                                                                     \
  }
    DDX_TILINGS16(DDX_T)
#undef DDX_T
    return false;
  }
  if (t.cy == 0 && t.kz > 0) {
    dim3 bl2(LX * LY);
    switch (mode) {
      case PURE_FORWARD:
      case RTM_FORWARD:
        hipLaunchKernelGGL(k_vel_lds<PURE_FORWARD>, g, bl2, 0, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R, t.kz);
        break;
      case RTM_BACKWARD:
        hipLaunchKernelGGL(k_vel_lds<RTM_BACKWARD>, g, bl2, 0, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R, t.kz);
        break;
      default:
        hipLaunchKernelGGL(k_vel_lds<RTM_REVERSE>, g, bl2, 0, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R, t.kz);
    }
    return true;
  }
  if (t.kz > 0) {
#define DDX_T(TX, TY, CY, CZ)                                                                      \
  if (t.tx == TX && t.ty == TY && t.cy == CY && t.kz == CZ) {                                      \
    vel_zb<TX, TY, CY, CZ, F32>(mode, g, st, d, starts, n, dx, dy, dz, sign, gnx, gny, gnz, R);     \
    return true;                                                                                  \
  }
    DDX_TILINGS(DDX_T)
#undef DDX_T
    return false;
  }
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
  return true;
}

bool launch_stress_phy(const Tiling &t, int mode, int fmt, const PhyDesc *d, const int *starts, int n, int nblk,
                       float dx, float dy, float dz, float dt, int gnx, int gny, int gnz, int R, hipStream_t st) {
  if (nblk <= 0) return true;
  if ((mode == FWI || fmt != STORE_F32) && !(t.kz > 0 && t.cy > 0)) return false;
  if (mode == FWI && (t.px != 1 || t.xl)) return false;
  dim3 g(nblk), bl(kBlock);
  if (t.xl == 2) {
    if (fmt != STORE_F32 || t.tx != 64 || t.cy != 1) return false;
#define DDX_T(TY, CZ)                                                                                  \
  if (t.ty == TY && t.kz == CZ) {                                                                      \
    const dim3 b2(64 * TY);                                                                            \
    switch (mode) {                                                                                    \
      case PURE_FORWARD:                                                                               \
        hipLaunchKernelGGL((k_stress_phy_yl<PURE_FORWARD, TY, CZ, F32>), g, b2, 0, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R); \
        break;                                                                                         \
      case RTM_FORWARD:                                                                                \
        hipLaunchKernelGGL((k_stress_phy_yl<RTM_FORWARD, TY, CZ, F32>), g, b2, 0, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R); \
        break;                                                                                         \
      case RTM_BACKWARD:                                                                               \
        hipLaunchKernelGGL((k_stress_phy_yl<RTM_BACKWARD, TY, CZ, F32>), g, b2, 0, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R); \
        break;                                                                                         \
      default:                                                                                         \
        hipLaunchKernelGGL((k_stress_phy_yl<RTM_REVERSE, TY, CZ, F32>), g, b2, 0, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R); \
    }                                                                                                  \
    return true;                                                                                      \
  }
    DDX_TILINGSY(DDX_T)
#undef DDX_T
    return false;
  }
  if (t.xl) {
    if (fmt != STORE_F32) return false;
#define DDX_T(TX, TY, CY, CZ)                                                                          \
  if (t.tx == TX && t.ty == TY && t.cy == CY && t.kz == CZ) {                                          \
    stress_zb<TX, TY, CY, CZ, F32, true>(mode, g, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R);  \
    return true;                                                                                      \
  }
    DDX_TILINGSX(DDX_T)
#undef DDX_T
    return false;
  }
  if (t.px == 2) {
#define DDX_T(TX, TY, CY, CZ)                                                                          \
  if (t.tx == TX && t.ty == TY && t.cy == CY && t.kz == CZ) {                                          \
    if (fmt == STORE_F16)                                                                              \
      stress_zb2<TX, TY, CY, CZ, F16>(mode, g, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R);    \
    else if (fmt == STORE_BF16)                                                                        \
      stress_zb2<TX, TY, CY, CZ, BF16>(mode, g, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R);   \
    else                                                                                               \
      stress_zb2<TX, TY, CY, CZ, F32>(mode, g, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R);    \
    return true;                                                                                      \
  }
    DDX_TILINGS2(DDX_T)
#undef DDX_T
    return false;
  }
  if (fmt == STORE_F16 || fmt == STORE_BF16) {
#define DDX_T(TX, TY, CY, CZ)                                                                          \
  if (t.tx == TX && t.ty == TY && t.cy == CY && t.kz == CZ) {                                          \
    if (fmt == STORE_F16)                                                                              \
      stress_zb<TX, TY, CY, CZ, F16>(mode, g, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R);     \
    else                                                                                               \
      stress_zb<TX, TY, CY, CZ, BF16>(mode, g, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R);    \
    return true;                                                                                      \
  }
    DDX_TILINGS16(DDX_T)
#undef DDX_T
    return false;
  }
  if (t.cy == 0 && t.kz > 0) {
    dim3 bl2(LX * LY);
    switch (mode) {
      case PURE_FORWARD:
        hipLaunchKernelGGL(k_stress_lds<PURE_FORWARD>, g, bl2, 0, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R, t.kz);
        break;
      case RTM_FORWARD:
        hipLaunchKernelGGL(k_stress_lds<RTM_FORWARD>, g, bl2, 0, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R, t.kz);
        break;
      case RTM_BACKWARD:
        hipLaunchKernelGGL(k_stress_lds<RTM_BACKWARD>, g, bl2, 0, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R, t.kz);
        break;
      default:
        hipLaunchKernelGGL(k_stress_lds<RTM_REVERSE>, g, bl2, 0, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R, t.kz);
    }
    return true;
  }
  if (t.kz > 0) {
#define DDX_T(TX, TY, CY, CZ)                                                                      \
  if (t.tx == TX && t.ty == TY && t.cy == CY && t.kz == CZ) {                                      \
    stress_zb<TX, TY, CY, CZ, F32>(mode, g, st, d, starts, n, dx, dy, dz, dt, gnx, gny, gnz, R);    \
    return true;                                                                                  \
  }
    DDX_TILINGS(DDX_T)
#undef DDX_T
    return false;
  }
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
  return true;
}

void launch_halo(int esz, const HaloDesc *d, const int *starts, int n, int nblk, hipStream_t st) {
  if (nblk <= 0) return;
  if (esz == 2) hipLaunchKernelGGL(k_halo<unsigned short>, dim3(nblk), dim3(kBlock), 0, st, d, starts, n);
  else hipLaunchKernelGGL(k_halo<float>, dim3(nblk), dim3(kBlock), 0, st, d, starts, n);
}

template <int C, class S>
static void vel_pml_launch(const SlabDesc *d, const int *starts, int n, int grid, float dx, float dy, float dz,
                           float rv, int grp, const PmlOrder &o, hipStream_t st) {
  hipLaunchKernelGGL((k_vel_pml<C, S>), dim3(grid), dim3(kBlock), 0, st, d, starts, n, dx, dy, dz, rv, grp, o);
}

// The grid of a CPML launch: the table's blocks, or 8 x the largest XCD
// share of the XCD-partitioned order.
static int pml_grid(int nblk, const PmlOrder *ord, PmlOrder &o) {
  if (ord && ord->pieces) {
    o = *ord;
    return o.grid;
  }
  o = PmlOrder{};
  return nblk;
}

template <int P>
static void vel_pml_t_launch(const SlabDesc *d, const int *starts, int n, int grid, float dx, float dy, float dz,
                             float, int grp, const PmlOrder &o, hipStream_t st) {
  hipLaunchKernelGGL((k_vel_pml_t<P>), dim3(grid), dim3(kBlock), 0, st, d, starts, n, dx, dy, dz, grp, o);
}

// The 16-bit formats run one cell per thread (cz is ignored for them; the
// tables of a 16-bit context are built for cz = 1).
void launch_vel_pml(int cz, int fmt, const SlabDesc *d, const int *starts, int n, int nblk, float dx, float dy,
                    float dz, float rv, int grp, const PmlOrder *ord, hipStream_t st) {
  if (nblk <= 0) return;
  PmlOrder o;
  const int grid = pml_grid(nblk, ord, o);
  if (grid <= 0) return;
  void (*f)(const SlabDesc *, const int *, int, int, float, float, float, float, int, const PmlOrder &,
            hipStream_t);
  if (fmt == STORE_F32 && pml_tiled(cz))
    f = cz == kPmlTiled + 2 ? vel_pml_t_launch<2> : cz == kPmlTiled + 4 ? vel_pml_t_launch<4> : vel_pml_t_launch<8>;
  else if (fmt == STORE_F16) f = vel_pml_launch<1, F16>;
  else if (fmt == STORE_BF16) f = vel_pml_launch<1, BF16>;
  else f = cz == 1 ? vel_pml_launch<1, F32> : cz == 2 ? vel_pml_launch<2, F32>
           : cz == 4 ? vel_pml_launch<4, F32> : vel_pml_launch<8, F32>;
  f(d, starts, n, grid, dx, dy, dz, rv, grp, o, st);
}

template <int W, int C, class S>
static void stress_pml_launch(const SlabDesc *d, const int *starts, int n, int grid, float dx, float dy, float dz,
                              float dt, int grp, const PmlOrder &o, hipStream_t st) {
  auto kern = k_stress_pml<W, C, S>;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, st, d, starts, n, dx, dy, dz, dt, grp, o);
}

template <int W, int P>
static void stress_pml_t_launch(const SlabDesc *d, const int *starts, int n, int grid, float dx, float dy, float dz,
                                float dt, int grp, const PmlOrder &o, hipStream_t st) {
  auto kern = k_stress_pml_t<W, P>;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, st, d, starts, n, dx, dy, dz, dt, grp, o);
}

// The 16-bit formats run one cell per thread (cz is ignored for them).
void launch_stress_pml(int cz, int with_sau, int fmt, const SlabDesc *d, const int *starts, int n, int nblk,
                       float dx, float dy, float dz, float dt, int grp, const PmlOrder *ord, hipStream_t st) {
  if (nblk <= 0) return;
  PmlOrder o;
  const int grid = pml_grid(nblk, ord, o);
  if (grid <= 0) return;
  void (*f)(const SlabDesc *, const int *, int, int, float, float, float, float, int, const PmlOrder &,
            hipStream_t);
  if (fmt == STORE_F32 && pml_tiled(cz)) {
    if (with_sau)
      f = cz == kPmlTiled + 2 ? stress_pml_t_launch<1, 2> : cz == kPmlTiled + 4 ? stress_pml_t_launch<1, 4>
          : stress_pml_t_launch<1, 8>;
    else
      f = cz == kPmlTiled + 2 ? stress_pml_t_launch<0, 2> : cz == kPmlTiled + 4 ? stress_pml_t_launch<0, 4>
          : stress_pml_t_launch<0, 8>;
  } else if (fmt == STORE_F16)
    f = with_sau ? stress_pml_launch<1, 1, F16> : stress_pml_launch<0, 1, F16>;
  else if (fmt == STORE_BF16)
    f = with_sau ? stress_pml_launch<1, 1, BF16> : stress_pml_launch<0, 1, BF16>;
  else if (with_sau)
    f = cz == 1 ? stress_pml_launch<1, 1, F32> : cz == 2 ? stress_pml_launch<1, 2, F32>
        : cz == 4 ? stress_pml_launch<1, 4, F32> : stress_pml_launch<1, 8, F32>;
  else
    f = cz == 1 ? stress_pml_launch<0, 1, F32> : cz == 2 ? stress_pml_launch<0, 2, F32>
        : cz == 4 ? stress_pml_launch<0, 4, F32> : stress_pml_launch<0, 8, F32>;
  f(d, starts, n, grid, dx, dy, dz, dt, grp, o, st);
}

void launch_copy(int dir, int esz, const CopyDesc *d, const int *starts, int n, int nblk, void *buf, hipStream_t st) {
  if (nblk <= 0) return;
  if (esz == 2) {
    if (dir > 0) hipLaunchKernelGGL((k_copy<1, unsigned short>), dim3(nblk), dim3(kBlock), 0, st, d, starts, n, buf);
    else hipLaunchKernelGGL((k_copy<-1, unsigned short>), dim3(nblk), dim3(kBlock), 0, st, d, starts, n, buf);
  } else {
    if (dir > 0) hipLaunchKernelGGL((k_copy<1, float>), dim3(nblk), dim3(kBlock), 0, st, d, starts, n, buf);
    else hipLaunchKernelGGL((k_copy<-1, float>), dim3(nblk), dim3(kBlock), 0, st, d, starts, n, buf);
  }
}

void launch_copy_conv(int dir, int fmt, const CopyDesc *d, const int *starts, int n, int nblk, void *buf, float sc_v,
                      float sc_s, hipStream_t st) {
  if (nblk <= 0) return;
  const dim3 g(nblk), b(kBlock);
  if (fmt == STORE_F16) {
    if (dir > 0) hipLaunchKernelGGL((k_copy_conv<1, F16>), g, b, 0, st, d, starts, n, buf, sc_v, sc_s);
    else hipLaunchKernelGGL((k_copy_conv<-1, F16>), g, b, 0, st, d, starts, n, buf, sc_v, sc_s);
  } else {
    if (dir > 0) hipLaunchKernelGGL((k_copy_conv<1, BF16>), g, b, 0, st, d, starts, n, buf, sc_v, sc_s);
    else hipLaunchKernelGGL((k_copy_conv<-1, BF16>), g, b, 0, st, d, starts, n, buf, sc_v, sc_s);
  }
}

void launch_ring_max(const CopyDesc *d, const int *starts, int n, int nblk, unsigned *maxbits, hipStream_t st) {
  if (nblk <= 0) return;
  hipLaunchKernelGGL(k_ring_max, dim3(nblk), dim3(kBlock), 0, st, d, starts, n, maxbits);
}

void launch_ring_exp(unsigned *maxbits, int *exps, int nf, hipStream_t st) {
  hipLaunchKernelGGL(k_ring_exp, dim3(1), dim3(64), 0, st, maxbits, exps, nf);
}

void launch_copy_dyn(int dir, int fmt, const CopyDesc *d, const int *starts, int n, int nblk, void *buf,
                     const int *exps, hipStream_t st) {
  if (nblk <= 0) return;
  const dim3 g(nblk), b(kBlock);
  if (fmt == STORE_F16) {
    if (dir > 0) hipLaunchKernelGGL((k_copy_dyn<1, F16>), g, b, 0, st, d, starts, n, buf, exps);
    else hipLaunchKernelGGL((k_copy_dyn<-1, F16>), g, b, 0, st, d, starts, n, buf, exps);
  } else {
    if (dir > 0) hipLaunchKernelGGL((k_copy_dyn<1, BF16>), g, b, 0, st, d, starts, n, buf, exps);
    else hipLaunchKernelGGL((k_copy_dyn<-1, BF16>), g, b, 0, st, d, starts, n, buf, exps);
  }
}

void launch_ring_bfp(int dir, const CopyDesc *d, const int *starts, int n, int nblk, void *buf, signed char *bexp,
                     hipStream_t st) {
  if (nblk <= 0) return;
  if (dir > 0) hipLaunchKernelGGL(k_save_bfp, dim3(nblk), dim3(kBlock), 0, st, d, starts, n, buf, bexp);
  else hipLaunchKernelGGL(k_inject_bfp, dim3(nblk), dim3(kBlock), 0, st, d, starts, n, (const void *)buf,
                          (const signed char *)bexp);
}

void launch_source(int fmt, void *sxx, void *syy, void *szz, long off, const float *wavelet, int it, float sign,
                   int src_mode, float ss, const int *itp, hipStream_t st) {
  if (fmt == STORE_F16)
    hipLaunchKernelGGL(k_source<F16>, dim3(1), dim3(1), 0, st, sxx, syy, szz, off, wavelet, it, sign, src_mode, ss,
                       itp);
  else if (fmt == STORE_BF16)
    hipLaunchKernelGGL(k_source<BF16>, dim3(1), dim3(1), 0, st, sxx, syy, szz, off, wavelet, it, sign, src_mode, ss,
                       itp);
  else
    hipLaunchKernelGGL(k_source<F32>, dim3(1), dim3(1), 0, st, sxx, syy, szz, off, wavelet, it, sign, src_mode, ss,
                       itp);
}

// *p += delta (one thread): the step counter of graph-replayed time steps;
// k_set_int sets it (a kernel, not a memset, so it may sit next to graphs)
__global__ void k_advance(int *p, int delta) { *p += delta; }
__global__ void k_set_int(int *p, int v) { *p = v; }
void launch_advance(int *p, int delta, hipStream_t st) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(1), 0, st, p, delta);
}
void launch_set_int(int *p, int v, hipStream_t st) { hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, p, v); }

static int recv_grid(int nrec) {
  int g = (nrec + kBlock - 1) / kBlock;
  return g > 1024 ? 1024 : g;
}

void launch_source_batch(int fmt, const MemberIO *io, int n, int gs, const int *gsp, hipStream_t st,
                         float sign) {
  if (n <= 0) return;
  const dim3 g((n + 63) / 64), b(64);
  if (fmt == STORE_F16) hipLaunchKernelGGL(k_source_batch<F16>, g, b, 0, st, io, n, gs, gsp, sign);
  else if (fmt == STORE_BF16) hipLaunchKernelGGL(k_source_batch<BF16>, g, b, 0, st, io, n, gs, gsp, sign);
  else hipLaunchKernelGGL(k_source_batch<F32>, g, b, 0, st, io, n, gs, gsp, sign);
}

void launch_inject_batch(int fmt, const MemberIO *io, const int *starts, int n, int nblk, int gs, hipStream_t st) {
  if (n <= 0 || nblk <= 0) return;
  const dim3 g(nblk), b(kBlock);
  if (fmt == STORE_F16) hipLaunchKernelGGL(k_inject_batch<F16>, g, b, 0, st, io, starts, n, gs);
  else if (fmt == STORE_BF16) hipLaunchKernelGGL(k_inject_batch<BF16>, g, b, 0, st, io, starts, n, gs);
  else hipLaunchKernelGGL(k_inject_batch<F32>, g, b, 0, st, io, starts, n, gs);
}

void launch_recv_batch(int fmt, const MemberIO *io, const int *starts, int n, int nblk, int gs, const int *gsp,
                       hipStream_t st) {
  if (n <= 0 || nblk <= 0) return;
  const dim3 g(nblk), b(kBlock);
  if (fmt == STORE_F16) hipLaunchKernelGGL(k_recv_batch<F16>, g, b, 0, st, io, starts, n, gs, gsp);
  else if (fmt == STORE_BF16) hipLaunchKernelGGL(k_recv_batch<BF16>, g, b, 0, st, io, starts, n, gs, gsp);
  else hipLaunchKernelGGL(k_recv_batch<F32>, g, b, 0, st, io, starts, n, gs, gsp);
}

void launch_recv_sample(int fmt, void *const *vptr, const int *rsub, const long *roff, int nrec, int it, int ntime,
                        float *seis, float inv_sv, const int *itp, hipStream_t st) {
  if (nrec <= 0) return;
  const dim3 g(recv_grid(nrec)), b(kBlock);
  if (fmt == STORE_F16)
    hipLaunchKernelGGL(k_recv_sample<F16>, g, b, 0, st, vptr, rsub, roff, nrec, it, ntime, seis, inv_sv, itp);
  else if (fmt == STORE_BF16)
    hipLaunchKernelGGL(k_recv_sample<BF16>, g, b, 0, st, vptr, rsub, roff, nrec, it, ntime, seis, inv_sv, itp);
  else
    hipLaunchKernelGGL(k_recv_sample<F32>, g, b, 0, st, vptr, rsub, roff, nrec, it, ntime, seis, inv_sv, itp);
}

void launch_recv_inject(int fmt, void *const *vptr, const int *rsub, const long *roff, int nrec, int it, int ntime,
                        const float *seis, float sv, hipStream_t st) {
  if (nrec <= 0) return;
  const dim3 g(recv_grid(nrec)), b(kBlock);
  if (fmt == STORE_F16)
    hipLaunchKernelGGL(k_recv_inject<F16>, g, b, 0, st, vptr, rsub, roff, nrec, it, ntime, seis, sv);
  else if (fmt == STORE_BF16)
    hipLaunchKernelGGL(k_recv_inject<BF16>, g, b, 0, st, vptr, rsub, roff, nrec, it, ntime, seis, sv);
  else
    hipLaunchKernelGGL(k_recv_inject<F32>, g, b, 0, st, vptr, rsub, roff, nrec, it, ntime, seis, sv);
}

void launch_recv_residual(void *const *vptr, const int *rsub, const long *roff, int nrec, int it, int ntime,
                          const float *obs, const float *syn, hipStream_t st) {
  if (nrec <= 0) return;
  hipLaunchKernelGGL(k_recv_residual, dim3(recv_grid(nrec)), dim3(kBlock), 0, st, vptr, rsub, roff, nrec, it, ntime,
                     obs, syn);
}

void launch_image(int fp64, int fmt, const ImgDesc *d, const int *starts, int n, int nblk, float inv_sv,
                  hipStream_t st) {
  if (nblk <= 0) return;
  const dim3 g(nblk), b(kBlock);
  if (fmt == STORE_F16) {
    if (fp64) hipLaunchKernelGGL((k_image<double, F16>), g, b, 0, st, d, starts, n, inv_sv);
    else hipLaunchKernelGGL((k_image<float, F16>), g, b, 0, st, d, starts, n, inv_sv);
  } else if (fmt == STORE_BF16) {
    if (fp64) hipLaunchKernelGGL((k_image<double, BF16>), g, b, 0, st, d, starts, n, inv_sv);
    else hipLaunchKernelGGL((k_image<float, BF16>), g, b, 0, st, d, starts, n, inv_sv);
  } else {
    if (fp64) hipLaunchKernelGGL((k_image<double, F32>), g, b, 0, st, d, starts, n, inv_sv);
    else hipLaunchKernelGGL((k_image<float, F32>), g, b, 0, st, d, starts, n, inv_sv);
  }
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

// Non-finite check of one field: a code whose exponent bits are all ones is
// an infinity or a NaN in every storage format.  Grid-stride; any thread
// that finds one stores 1 (all such stores write the same value).
template <class S>
__global__ __launch_bounds__(kBlock) void k_nonfinite(const void *__restrict__ field, long n, int *flag) {
  int bad = 0;
  for (long i = (long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long)gridDim.x * kBlock) {
    if (std::is_same<S, F32>::value) {
      const unsigned u = ((const __attribute__((address_space(1))) unsigned *)field)[i];
      bad |= (u & 0x7f800000u) == 0x7f800000u;
    } else {
      const unsigned short u = ((const __attribute__((address_space(1))) unsigned short *)field)[i];
      const unsigned short m = std::is_same<S, F16>::value ? 0x7c00 : 0x7f80;
      bad |= (u & m) == m;
    }
  }
  if (bad) *flag = 1;
}

void launch_nonfinite(int fmt, const void *field, long n, int *flag, hipStream_t st) {
  if (n <= 0) return;
  const long blocks = std::min<long>((n + kBlock - 1) / kBlock, 2048);
  const dim3 g((unsigned)blocks), b(kBlock);
  if (fmt == STORE_F16) hipLaunchKernelGGL(k_nonfinite<F16>, g, b, 0, st, field, n, flag);
  else if (fmt == STORE_BF16) hipLaunchKernelGGL(k_nonfinite<BF16>, g, b, 0, st, field, n, flag);
  else hipLaunchKernelGGL(k_nonfinite<F32>, g, b, 0, st, field, n, flag);
}

}  // namespace ddx
