// =============================================================================
// CPU ORACLE: TEST INFRASTRUCTURE ONLY.
//
// This file restates, in plain C++ (templated on float / double), the
// algorithm of CalvinStudio/domain-decomposition-release for 3D elastic
// staggered-grid FDTD modelling and RTM.  It is the checker for the HIP
// engine: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it.  The product library (libddx) never links it.
//
// Formulation.  The reference splits the global extended grid into MPI
// subdomains whose edge kernels read neighbour halos at the full stencil
// order, so the decomposition does not change the arithmetic of any cell.
// The oracle therefore updates the GLOBAL extended grid directly and keeps
// the decomposition only where it changes outputs: which receivers a
// subdomain keeps, the per-subdomain image Laplacian border, and the order
// of the per-subdomain seismogram rows.
//
// Every function cites the reference file:line it follows (paths relative
// to the reference root; aliases):
//   PHY/  = src/jarvis_cpp/jarvis/.jarvis_2_app/.jarvis_3_seismology/
//           elastic_1_fdtd3d_model_mpi/fdtd3d_2_phy_mpi_module/
//   CPML/ = .../elastic_1_fdtd3d_model_mpi/fdtd3d_1_cpml_mpi_module/
//   BASE/ = .../elastic_1_fdtd3d_model_mpi/fdtd3d_0_base_module/
//   BLEND/= .../elastic_1_fdtd3d_model_mpi/fdtd3d_3_model_mpi_blend/
//   IMG/  = .../elastic_1_rtm3d_model_mpi/rtm3d_0_image_module/
//   RTM/  = .../elastic_1_rtm3d_model_mpi/rtm3d_1_model_mpi_blend/
//   FLD/  = src/jarvis_cpp/jarvis/.jarvis_0_meta/.jarvis_2_field/
//   GEO/  = src/jarvis_cpp/jarvis/.jarvis_1_util/.jarvis_2_geometry/
//
// Build with -ffp-contract=off so every float operation is a single IEEE
// rounding in the reference's order of operations.
// =============================================================================
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

static const double kPi = 3.14159265358979323846264338327950288;

// Planted mistakes.  The pins in tests/ switch one on (orc_plant_fault) to
// show that they fail on a broken oracle; 0, the default, is the faithful
// restatement.  Each names the term it breaks.
enum Fault {
  F_NONE = 0,
  F_INJ_TIME = 1,     // reverse receiver injection reads sample it - 1
  F_INJ_SIGN = 2,     // reverse receiver injection subtracts
  F_INJ_COMP = 3,     // reverse receiver injection: vx data into vy, vy into vz, vz into vx
  F_SRC_SZZ = 4,      // vertical-stress source adds into sxx
  F_SHEAR_LAMBDA = 5, // interior sxy update uses lambda for mu
  F_FD_COEF = 6,      // 8th-order coefficient c[1] off by 1%
  F_PML_VX_X = 7,     // CPML velocity: d sxx/dx damped with the full-grid a, b
  F_PML_VY_Y = 8,     // CPML velocity: d syy/dy damped with the full-grid a, b
  F_PML_VZ_Z = 9,     // CPML velocity: d szz/dz damped with the full-grid a, b
  F_PML_SXX_X = 10,   // CPML stress: d vx/dx damped with the half-grid a, b
  F_PML_SXY_X = 11,   // CPML stress: d vy/dx (sxy) damped with the full-grid a, b
  F_PML_VY_SXY_X = 12 // CPML velocity: d sxy/dx (vy) damped with the half-grid a, b
};
static int g_fault = F_NONE;
// true: every update takes the per-cell code, never the vectorized runs
static bool g_scalar = false;

// FLD/jarvis_3_device_field_meat.hpp:9-22 (get_gap_num): node distance with
// a +-0.125 truncation bias, evaluated in float.
static inline int gap_num(float l, float r, float d) {
  float v = 1.0f / d;
  int nl = (l > 0) ? (int)(l * v + 0.125f) : (int)(l * v - 0.125f);
  int nr = (r > 0) ? (int)(r * v + 0.125f) : (int)(r * v - 0.125f);
  return nr - nl;
}

// GEO/jarvis_device_geometry_bones.hpp:29-49 (cu_cal_diff_coeff): staggered
// first-derivative coefficients of half order `half`, computed in double.
static void fd_coef_double(int half, double *out) {
  for (int n = 0; n < half; n++) {
    double a1 = 1, a2 = 1, a3 = 1;
    for (int i = 0; i < half; i++) {
      if (i != n) a1 = a1 * (2 * i + 1) * (2 * i + 1);
      if (i < n) a2 = a2 * ((2 * n + 1) * (2 * n + 1) - (2 * i + 1) * (2 * i + 1));
      if (i > n) a3 = a3 * ((2 * i + 1) * (2 * i + 1) - (2 * n + 1) * (2 * n + 1));
    }
    a1 = a1 * std::pow(-1.0, n);
    out[n] = a1 / ((2 * n + 1) * a2 * a3);
  }
}

// ---------------------------------------------------------------------------
// 16-bit wavefield storage (the engine's FP16 / BF16 storage formats; no
// counterpart in the reference, which stores FP32).  Conversions round to
// nearest, ties to even, written from the IEEE 754 binary16 and bfloat16
// encodings; pinned against numpy's float16 and torch's bfloat16 casts.
// ---------------------------------------------------------------------------
static inline uint32_t f32_bits(float x) { uint32_t u; std::memcpy(&u, &x, 4); return u; }
static inline float bits_f32(uint32_t u) { float x; std::memcpy(&x, &u, 4); return x; }

static uint16_t f32_to_f16(float x) {
  const uint32_t u = f32_bits(x);
  const uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
  const uint32_t a = u & 0x7fffffffu;
  if (a >= 0x7f800000u) return sign | (a > 0x7f800000u ? 0x7e00 : 0x7c00);  // NaN, Inf
  const int e = (int)(a >> 23);                                               // biased float exponent
  if (e >= 143) return sign | 0x7c00;                                          // >= 2^16: overflow
  if (e >= 113) {                                                              // normal binary16
    uint32_t r = ((uint32_t)(e - 112) << 10) | ((a >> 13) & 0x3ffu);
    const uint32_t rem = a & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) r++;  // a carry may reach the exponent / Inf
    return sign | (uint16_t)r;
  }
  if (e < 102) return sign;  // below 2^-25: rounds to zero
  // subnormal binary16: value = m * 2^(e-150) = (m >> (126-e)) * 2^-24
  const uint32_t m = (a & 0x7fffffu) | 0x800000u;
  const int shift = 126 - e;  // 14..24
  uint32_t r = m >> shift;
  const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
  if (rem > half || (rem == half && (r & 1u))) r++;
  return sign | (uint16_t)r;
}
static float f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  if (e == 0) {
    const float v = (float)m * 5.9604644775390625e-08f;  // m * 2^-24, exact
    return bits_f32(f32_bits(v) | sign);
  }
  if (e == 31) return bits_f32(sign | 0x7f800000u | (m << 13));
  return bits_f32(sign | ((e + 112u) << 23) | (m << 13));
}
static uint16_t f32_to_bf16(float x) {
  uint32_t u = f32_bits(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float bf16_to_f32(uint16_t b) { return bits_f32((uint32_t)b << 16); }

enum Storage { STORE_F32 = 0, STORE_F16 = 1, STORE_BF16 = 2, STORE_BFP16 = 3 };

// Block-scaled 16-bit VOWR store (ring format STORE_BFP16; a storage option
// of the engine with no counterpart in the reference).  Runs of kBfpRun
// consecutive entries share the exponent s of their largest |value| m
// (m = f 2^s, 0.5 <= f < 1; s = -126 for m zero or subnormal).  An entry
// keeps u = |v| 2^-s in [0, 1): for u >= 2^-7 as (4096 + M) 2^(E - 20),
// E = 1..7 its binade and M a 12-bit mantissa, rounded to nearest (ties to
// even; a mantissa that rounds up moves to the next binade); below 2^-7 as
// the nearest multiple of 2^-19.  The largest code, 32767, is 8191 2^-13:
// u that rounds up to 1 keeps it.
constexpr int kBfpRun = 16;
static int bfp_exponent(float m) {
  if (!(m >= 1.17549435e-38f) || !(m <= 3.4028235e38f)) return -126;
  int e = 0;
  std::frexp(m, &e);
  return std::max(e, -126);
}
static float bfp_round(float v, int s) {
  const float u = std::ldexp(std::fabs(v), -s);
  int c;
  if (u >= 0.0078125f) {
    int ex = 0;
    std::frexp(u, &ex);
    c = 4096 * (ex + 6) + (int)std::rint(std::ldexp(u, 13 - ex));
  } else {
    c = (int)std::rint(std::ldexp(u, 19));
  }
  c = std::min(c, 32767);
  const float r = c < 4096 ? std::ldexp((float)c, -19) : std::ldexp((float)(4096 + (c & 4095)), (c >> 12) - 20);
  return v < 0 ? -std::ldexp(r, s) : std::ldexp(r, s);
}

// Value a field holds after a store: x scaled by the power of two s, rounded
// to the storage format, scaled back (every step exact except the rounding).
static inline float quant(int fmt, float x, float s) {
  if (fmt == STORE_F16) return f16_to_f32(f32_to_f16(x * s)) / s;
  if (fmt == STORE_BF16) return bf16_to_f32(f32_to_bf16(x * s)) / s;
  return x;
}
static inline double quant(int, double x, float) { return x; }  // the FP64 oracle stores exactly

// Field slots (order is the oracle's own; the engine uses the same ids).
enum Field {
  VX = 0, VY, VZ, SXX, SYY, SZZ, SXY, SXZ, SYZ, SAU, VXP, VYP, VZP,
  // FWI accumulators (BASE/fdtd3d_base_2_wavefield_bones.hpp:19-22, fwiWave)
  UX, UY, UZ, VXRT, VYRT, VZRT, NFIELD
};
// Propagation modes (BASE/fdtd3d_base_1_interface.hpp:19-26 SimulateType).
enum Mode { PURE_FORWARD = 0, RTM_FORWARD = 1, RTM_BACKWARD = 2, RTM_REVERSE = 3, FWI = 4 };
// CPML memory variables, named by the derivative they damp.
enum Psi {
  P_SXX_X = 0, P_SXY_Y, P_SXZ_Z,  // vx update
  P_SXY_X, P_SYY_Y, P_SYZ_Z,      // vy update
  P_SXZ_X, P_SYZ_Y, P_SZZ_Z,      // vz update
  P_VX_X, P_VY_Y, P_VZ_Z,         // normal stresses
  P_VY_X, P_VX_Y,                 // sxy
  P_VZ_X, P_VX_Z,                 // sxz
  P_VZ_Y, P_VY_Z,                 // syz
  NPSI
};

// VOWR ring widths, order (top, bottom, front, back, right, left):
// BASE/fdtd3d_base_5_wave_recon_meat.hpp:19-30.  Returned for field f and ring
// width R.  Sides: 0 top (low z), 1 bottom (high z), 2 front (low y),
// 3 back (high y), 4 right (low x), 5 left (high x).
static int ring_width(int f, int side, int R) {
  static const int tab[10][6] = {
      // entries: 1 means R, 0 means R-1, -1 means none
      {0, 1, 0, 1, 1, 0},     // vx
      {0, 1, 1, 0, 0, 1},     // vy
      {1, 0, 0, 1, 0, 1},     // vz
      {-1, -1, -1, -1, 0, 1}, // sxx
      {-1, -1, 0, 1, -1, -1}, // syy
      {0, 1, -1, -1, -1, -1}, // szz
      {-1, -1, 1, 0, 1, 0},   // sxy
      {1, 0, -1, -1, 1, 0},   // sxz
      {1, 0, 1, 0, -1, -1},   // syz
      {0, 1, 0, 1, 0, 1},     // sau
  };
  if (f < 0 || f > SAU) return 0;
  int t = tab[f][side];
  if (t < 0) return 0;
  return t ? R : R - 1;
}

template <class T>
struct Wave {
  std::vector<T> f[NFIELD];
  std::vector<T> p[NPSI];
  void alloc(size_t n) {
    for (auto &a : f) a.assign(n, T(0));
    for (auto &a : p) a.assign(n, T(0));
  }
  void zero() {
    for (auto &a : f) std::fill(a.begin(), a.end(), T(0));
    for (auto &a : p) std::fill(a.begin(), a.end(), T(0));
  }
};

struct Desc {
  int nx, ny, nz;        // global physical grid
  float dx, dy, dz;
  int pml;               // margin per side (pml_num), includes the 2 rigid layers
  float dt, fm;
  int ntime;
  int rec_width;         // VOWR ring width R
};

template <class T>
struct Sim {
  Desc d;
  int NX, NY, NZ, m;
  size_t NXY, N;
  T dx, dy, dz, dt;
  std::vector<T> lam, mu, rho, vp;
  // CPML coefficients, dense on the extended grid; only cells inside the
  // damping range of an axis are ever read for that axis.
  std::vector<T> ax, bx, axh, bxh, ay, by, ayh, byh, az, bz, azh, bzh;
  T ctab[5][4];  // ctab[h][ir]: coefficients of half order h (1..4)
  T c4[2];       // pml_fdorder_half = 2
  Wave<T> w[2];  // w[0]: forward/backward set, w[1]: reverse set
  std::vector<T> img[5];  // Ip, Ipp, Ips, Isp, Iss on the global gridphy
  // Storage format of the wavefields and CPML memory variables, with the
  // power-of-two scales of the velocity class (vx..vz, vxp..vzp and the
  // stress-update memory variables) and of the stress class (sxx..syz, sau
  // and the velocity-update memory variables).
  int fmt = STORE_F32;
  float sv = 1.f, ss = 1.f;
  T qv(T x) const { return quant(fmt, x, sv); }
  T qs(T x) const { return quant(fmt, x, ss); }
  // VOWR boundary store in a 16-bit format over FP32 wavefields: values are
  // quantized when saved (class scales rsv, rss)
  int ring_fmt = STORE_F32;
  float rsv = 1.f, rss = 1.f;
  // ring_dyn: each saved slice of each ring field takes its own power-of-two
  // scale 2^e, e = 15 - the binary exponent of the field's largest |value| in
  // the slice (frexp), clamped to [-100, 100], instead of the class scales
  bool ring_dyn = false;
  // STORE_BFP16 store: the subdomains' first gridphy index along x, y, z
  // (global gridphy coordinates, ascending, starting at 0): a ring box is
  // stored per subdomain part, its runs counted from each part's origin
  std::vector<int> rcut[3];

  size_t idx(int i, int j, int k) const { return (size_t)i + (size_t)NX * ((size_t)j + (size_t)NY * k); }
  bool in_phy(int i, int j, int k) const {
    return i >= m && i < m + d.nx && j >= m && j < m + d.ny && k >= m && k < m + d.nz;
  }
  static bool in_pml_axis(int i, int n, int mm) { return (i >= 2 && i < mm) || (i >= n - mm && i < n - 2); }
};

// BASE/fdtd3d_base_3_geomodel_meat.hpp:100-364 (extend_model): nearest-edge
// replication of the physical model into the margin (all 26 regions reduce to
// clamping each index into the physical range).
template <class T>
static void extend(const Sim<T> &s, const float *phy, std::vector<float> &ext) {
  ext.assign(s.N, 0.f);
  for (int k = 0; k < s.NZ; k++)
    for (int j = 0; j < s.NY; j++)
      for (int i = 0; i < s.NX; i++) {
        int pi = std::min(std::max(i - s.m, 0), s.d.nx - 1);
        int pj = std::min(std::max(j - s.m, 0), s.d.ny - 1);
        int pk = std::min(std::max(k - s.m, 0), s.d.nz - 1);
        ext[s.idx(i, j, k)] = phy[(size_t)pi + (size_t)s.d.nx * ((size_t)pj + (size_t)s.d.ny * pk)];
      }
}

// CPML/fdtd3d_cpml_mpi_1_host_meat.hpp:147-620 (calculate_damp_coeff).
// One axis profile at extended index i of an axis with N cells and margin m.
// Low-side slabs use the integer distance for the full grid and -0.5 for the
// half grid; high-side slabs the reverse (:186-189 vs :304-307).
template <class T>
static void cpml_coef_axis(const Sim<T> &s, float vpv, int i, int N, float h,
                           T &a, T &b, T &ah, T &bh) {
  const int mm = s.m;
  a = 0; b = 1; ah = 0; bh = 1;
  if (!Sim<T>::in_pml_axis(i, N, mm)) return;
  if (sizeof(T) == sizeof(float)) {
    const float R = 0.001f, DAMP = 7.0f;
    const float L = float(mm - 2);
    const float amax = (float)(kPi * (double)s.d.fm);
    float temp = std::log10(1.0f / R) * DAMP * vpv / (2.0f * L * h);
    float df, dh, alf, alh;
    if (i < mm) {
      float rf = (mm - i) / L, rh = (mm - i - 0.5f) / L;
      df = (float)(temp * std::pow((double)rf, 2));
      alf = amax * (1.0f - rf);
      dh = (float)(temp * std::pow((double)rh, 2));
      alh = amax * (1.0f - rh);
    } else {
      float rf = (i - (N - mm - 1) - 0.5f) / L, rh = (i - (N - mm - 1)) / L;
      df = (float)(temp * std::pow((double)rf, 2));
      alf = amax * (1.0f - rf);
      dh = (float)(temp * std::pow((double)rh, 2));
      alh = amax * (1.0f - rh);
    }
    if (alf < 0) alf = 0;
    if (alh < 0) alh = 0;
    float dtf = (float)s.d.dt;
    float bf = 0, bhf = 0, af = 0, ahf = 0;
    if (df >= 0) {
      bf = std::exp(-(df + alf) * dtf);
      bhf = std::exp(-(dh + alh) * dtf);
    }
    if (std::fabs(df) >= 1e-8) {
      af = df * (bf - 1.0f) / (df + alf);
      ahf = dh * (bhf - 1.0f) / (dh + alh);
    }
    a = af; b = bf; ah = ahf; bh = bhf;
  } else {
    const double R = 0.001, DAMP = 7.0;
    const double L = double(mm - 2);
    const double amax = kPi * (double)s.d.fm;
    double temp = std::log10(1.0 / R) * DAMP * (double)vpv / (2.0 * L * (double)h);
    double rf, rh;
    if (i < mm) { rf = (mm - i) / L; rh = (mm - i - 0.5) / L; }
    else { rf = (i - (N - mm - 1) - 0.5) / L; rh = (i - (N - mm - 1)) / L; }
    double df = temp * rf * rf, dh = temp * rh * rh;
    double alf = std::max(0.0, amax * (1.0 - rf)), alh = std::max(0.0, amax * (1.0 - rh));
    double dtd = (double)s.d.dt;
    double bf = std::exp(-(df + alf) * dtd), bhf = std::exp(-(dh + alh) * dtd);
    double af = 0, ahf = 0;
    if (std::fabs(df) >= 1e-8) { af = df * (bf - 1.0) / (df + alf); ahf = dh * (bhf - 1.0) / (dh + alh); }
    a = (T)af; b = (T)bf; ah = (T)ahf; bh = (T)bhf;
  }
}

template <class T>
static void init_sim(Sim<T> &s, const Desc &d, const float *vp, const float *vs, const float *rh) {
  s.d = d;
  s.m = d.pml;
  s.NX = d.nx + 2 * s.m;
  s.NY = d.ny + 2 * s.m;
  s.NZ = d.nz + 2 * s.m;
  s.NXY = (size_t)s.NX * s.NY;
  s.N = s.NXY * s.NZ;
  s.dx = (T)d.dx; s.dy = (T)d.dy; s.dz = (T)d.dz; s.dt = (T)d.dt;
  std::vector<float> evp, evs, erh;
  extend(s, vp, evp);
  extend(s, vs, evs);
  extend(s, rh, erh);
  s.lam.resize(s.N); s.mu.resize(s.N); s.rho.resize(s.N); s.vp.resize(s.N);
  // BASE/fdtd3d_base_3_geomodel_meat.hpp:456-460: lambda and mu in float.
  for (size_t i = 0; i < s.N; i++) {
    if (sizeof(T) == sizeof(float)) {
      s.lam[i] = erh[i] * (evp[i] * evp[i] - 2.0f * evs[i] * evs[i]);
      s.mu[i] = erh[i] * (evs[i] * evs[i]);
    } else {
      double p = evp[i], q = evs[i], r = erh[i];
      s.lam[i] = (T)(r * (p * p - 2.0 * q * q));
      s.mu[i] = (T)(r * (q * q));
    }
    s.rho[i] = (T)erh[i];
    s.vp[i] = (T)evp[i];
  }
  // coefficient tables
  for (int h = 1; h <= 4; h++) {
    double c[4] = {0, 0, 0, 0};
    fd_coef_double(h, c);
    for (int r = 0; r < 4; r++) s.ctab[h][r] = (T)c[r];
  }
  if (g_fault == F_FD_COEF) s.ctab[4][1] = s.ctab[4][1] * T(1.01);
  for (int r = 0; r < 4; r++) s.ctab[0][r] = 0;
  s.c4[0] = s.ctab[2][0];
  s.c4[1] = s.ctab[2][1];
  // CPML coefficients use the extended vp of each cell
  // (CPML/fdtd3d_cpml_mpi_1_host_meat.hpp:186 `model.vp(i, j, k)`).
  std::vector<T> *arrs[12] = {&s.ax, &s.bx, &s.axh, &s.bxh, &s.ay, &s.by, &s.ayh, &s.byh, &s.az, &s.bz, &s.azh, &s.bzh};
  for (auto *a : arrs) a->assign(s.N, T(0));
  for (int k = 0; k < s.NZ; k++)
    for (int j = 0; j < s.NY; j++)
      for (int i = 0; i < s.NX; i++) {
        size_t id = s.idx(i, j, k);
        float v = evp[id];
        cpml_coef_axis(s, v, i, s.NX, d.dx, s.ax[id], s.bx[id], s.axh[id], s.bxh[id]);
        cpml_coef_axis(s, v, j, s.NY, d.dy, s.ay[id], s.by[id], s.ayh[id], s.byh[id]);
        cpml_coef_axis(s, v, k, s.NZ, d.dz, s.az[id], s.bz[id], s.azh[id], s.bzh[id]);
      }
  s.w[0].alloc(s.N);
  s.w[1].alloc(s.N);
  size_t np = (size_t)d.nx * d.ny * d.nz;
  for (auto &a : s.img) a.assign(np, T(0));
}

// Stencil taps.  Dp: forward-staggered  sum c[r](f[i+(r+1)s] - f[i-r s]);
// Dm: backward-staggered sum c[r](f[i+r s] - f[i-(r+1)s]); accumulated from
// r = 0 upward as in PHY/fdtd3d_phy_mpi_2_vel_kernel.hpp:26-50.
template <class T>
static inline T Dp(const T *f, size_t i, size_t s, const T *c, int ord) {
  T acc = 0;
  for (int r = 0; r < ord; r++) acc += c[r] * (f[i + (size_t)(r + 1) * s] - f[i - (size_t)r * s]);
  return acc;
}
template <class T>
static inline T Dm(const T *f, size_t i, size_t s, const T *c, int ord) {
  T acc = 0;
  for (int r = 0; r < ord; r++) acc += c[r] * (f[i + (size_t)r * s] - f[i - (size_t)(r + 1) * s]);
  return acc;
}

// VOWR variable order (PHY/fdtd3d_phy_mpi_2_vel_kernel.hpp:132-152, :240-260):
// in backward mode, a cell at distance dd < phy_fdorder_half from the global
// physical boundary along an axis uses half order max(R, dd+1) on that axis.
static inline int vowr_order(int p, int n, int R) {
  int dd = std::min(p, n - 1 - p);
  if (dd >= 4) return 4;
  return std::max(R, dd + 1);
}

// Full-order (half order 4) runs of one x row, written so the compiler can
// vectorize them across i.  Per cell they perform exactly the operations of
// Dp/Dm and of the per-cell code below, in the same order (accumulator from
// 0, r = 0 upward, the same parenthesization), so results are identical;
// they serve rows whose three orders are all 4 and storage without
// quantization.
template <class T>
__attribute__((always_inline)) static inline T Dp4(const T *__restrict f, size_t i, size_t s, T c0, T c1, T c2, T c3) {
  T acc = 0;
  acc += c0 * (f[i + s] - f[i]);
  acc += c1 * (f[i + 2 * s] - f[i - s]);
  acc += c2 * (f[i + 3 * s] - f[i - 2 * s]);
  acc += c3 * (f[i + 4 * s] - f[i - 3 * s]);
  return acc;
}
template <class T>
__attribute__((always_inline)) static inline T Dm4(const T *__restrict f, size_t i, size_t s, T c0, T c1, T c2, T c3) {
  T acc = 0;
  acc += c0 * (f[i] - f[i - s]);
  acc += c1 * (f[i + s] - f[i - 2 * s]);
  acc += c2 * (f[i + 2 * s] - f[i - 3 * s]);
  acc += c3 * (f[i + 3 * s] - f[i - 4 * s]);
  return acc;
}

template <class T, bool RTMP>
static void vel_run4(Sim<T> &s, Wave<T> &w, T dt, size_t id0, int n) {
  T *__restrict vx = w.f[VX].data(), *__restrict vy = w.f[VY].data(), *__restrict vz = w.f[VZ].data();
  const T *__restrict sxx = w.f[SXX].data(), *__restrict syy = w.f[SYY].data(), *__restrict szz = w.f[SZZ].data();
  const T *__restrict sxy = w.f[SXY].data(), *__restrict sxz = w.f[SXZ].data(), *__restrict syz = w.f[SYZ].data();
  const T *__restrict sau = w.f[SAU].data();
  T *__restrict vxp = w.f[VXP].data(), *__restrict vyp = w.f[VYP].data(), *__restrict vzp = w.f[VZP].data();
  const T *__restrict rho = s.rho.data();
  const size_t SX = 1, SY = s.NX, SZ = s.NXY;
  const T c0 = s.ctab[4][0], c1 = s.ctab[4][1], c2 = s.ctab[4][2], c3 = s.ctab[4][3];
  const T hx = s.dx, hy = s.dy, hz = s.dz;
#pragma omp simd
  for (int q = 0; q < n; q++) {
    const size_t id = id0 + (size_t)q;
    T dtr = dt / rho[id];
    T a = dtr / hx, b = dtr / hy, c = dtr / hz;
    T sxx_x = Dp4(sxx, id, SX, c0, c1, c2, c3), syy_y = Dp4(syy, id, SY, c0, c1, c2, c3);
    T szz_z = Dp4(szz, id, SZ, c0, c1, c2, c3);
    T syx_x = Dm4(sxy, id, SX, c0, c1, c2, c3), sxy_y = Dm4(sxy, id, SY, c0, c1, c2, c3);
    T sxz_z = Dm4(sxz, id, SZ, c0, c1, c2, c3), szx_x = Dm4(sxz, id, SX, c0, c1, c2, c3);
    T szy_y = Dm4(syz, id, SY, c0, c1, c2, c3), syz_z = Dm4(syz, id, SZ, c0, c1, c2, c3);
    vx[id] = vx[id] + (a * sxx_x + b * sxy_y + c * sxz_z);
    vy[id] = vy[id] + (a * syx_x + b * syy_y + c * syz_z);
    vz[id] = vz[id] + (a * szx_x + b * szy_y + c * szz_z);
    if (RTMP) {
      T sx = Dp4(sau, id, SX, c0, c1, c2, c3), sy = Dp4(sau, id, SY, c0, c1, c2, c3);
      T sz = Dp4(sau, id, SZ, c0, c1, c2, c3);
      vxp[id] = vxp[id] + a * sx;
      vyp[id] = vyp[id] + b * sy;
      vzp[id] = vzp[id] + c * sz;
    }
  }
}

template <class T, bool WITH_SAU>
static void stress_run4(Sim<T> &s, Wave<T> &w, T dt, size_t id0, int n) {
  const T *__restrict vx = w.f[VX].data(), *__restrict vy = w.f[VY].data(), *__restrict vz = w.f[VZ].data();
  T *__restrict sxx = w.f[SXX].data(), *__restrict syy = w.f[SYY].data(), *__restrict szz = w.f[SZZ].data();
  T *__restrict sxy = w.f[SXY].data(), *__restrict sxz = w.f[SXZ].data(), *__restrict syz = w.f[SYZ].data();
  T *__restrict sau = w.f[SAU].data();
  const T *__restrict lam = s.lam.data(), *__restrict muv = s.mu.data();
  const size_t SX = 1, SY = s.NX, SZ = s.NXY;
  const T c0 = s.ctab[4][0], c1 = s.ctab[4][1], c2 = s.ctab[4][2], c3 = s.ctab[4][3];
  const T a = dt / s.dx, b = dt / s.dy, c = dt / s.dz;
#pragma omp simd
  for (int q = 0; q < n; q++) {
    const size_t id = id0 + (size_t)q;
    T lm = lam[id], mu = muv[id];
    T l2m = lm + T(2) * mu;
    T vx_x = Dm4(vx, id, SX, c0, c1, c2, c3), vx_y = Dp4(vx, id, SY, c0, c1, c2, c3);
    T vx_z = Dp4(vx, id, SZ, c0, c1, c2, c3);
    T vy_x = Dp4(vy, id, SX, c0, c1, c2, c3), vy_y = Dm4(vy, id, SY, c0, c1, c2, c3);
    T vy_z = Dp4(vy, id, SZ, c0, c1, c2, c3);
    T vz_x = Dp4(vz, id, SX, c0, c1, c2, c3), vz_y = Dp4(vz, id, SY, c0, c1, c2, c3);
    T vz_z = Dm4(vz, id, SZ, c0, c1, c2, c3);
    sxx[id] = sxx[id] + (a * l2m * vx_x + b * lm * vy_y + c * lm * vz_z);
    syy[id] = syy[id] + (a * lm * vx_x + b * l2m * vy_y + c * lm * vz_z);
    szz[id] = szz[id] + (a * lm * vx_x + b * lm * vy_y + c * l2m * vz_z);
    sxy[id] = sxy[id] + (a * mu * vy_x + b * mu * vx_y);
    sxz[id] = sxz[id] + (a * mu * vz_x + c * mu * vx_z);
    syz[id] = syz[id] + (b * mu * vz_y + c * mu * vy_z);
    if (WITH_SAU) sau[id] = sau[id] + (a * l2m * vx_x + b * l2m * vy_y + c * l2m * vz_z);
  }
}

// PHY/fdtd3d_phy_mpi_2_vel_kernel.hpp:7-80 (inside) and the edge kernels
// :82-2087; BWD variable order as above.
template <class T>
static void vel_phy(Sim<T> &s, Wave<T> &w, int mode, T dt, int R) {
  const bool bwd = mode == RTM_BACKWARD;
  const bool rtm_p = mode == RTM_BACKWARD || mode == RTM_REVERSE;
  T *vx = w.f[VX].data(), *vy = w.f[VY].data(), *vz = w.f[VZ].data();
  const T *sxx = w.f[SXX].data(), *syy = w.f[SYY].data(), *szz = w.f[SZZ].data();
  const T *sxy = w.f[SXY].data(), *sxz = w.f[SXZ].data(), *syz = w.f[SYZ].data();
  const T *sau = w.f[SAU].data();
  T *vxp = w.f[VXP].data(), *vyp = w.f[VYP].data(), *vzp = w.f[VZP].data();
  T *ux = w.f[UX].data(), *uy = w.f[UY].data(), *uz = w.f[UZ].data();
  T *vxrt = w.f[VXRT].data(), *vyrt = w.f[VYRT].data(), *vzrt = w.f[VZRT].data();
  const bool fwi = mode == FWI;
  const bool fast = s.fmt == STORE_F32 && !fwi && s.d.nx >= 8 && !g_scalar;
  const size_t SX = 1, SY = s.NX, SZ = s.NXY;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = s.m; k < s.m + s.d.nz; k++)
    for (int j = s.m; j < s.m + s.d.ny; j++) {
      int oz = bwd ? vowr_order(k - s.m, s.d.nz, R) : 4;
      int oy = bwd ? vowr_order(j - s.m, s.d.ny, R) : 4;
      const T *cz = s.ctab[oz], *cy = s.ctab[oy];
      // [f0, f1): the row's full-order run, updated by vel_run4
      int f0 = s.m + s.d.nx, f1 = f0;
      if (fast && oy == 4 && oz == 4) {
        f0 = s.m + (bwd ? 4 : 0);
        f1 = s.m + s.d.nx - (bwd ? 4 : 0);
        if (rtm_p) vel_run4<T, true>(s, w, dt, s.idx(f0, j, k), f1 - f0);
        else vel_run4<T, false>(s, w, dt, s.idx(f0, j, k), f1 - f0);
      }
      for (int i = s.m; i < s.m + s.d.nx; i++) {
        if (i == f0) { i = f1 - 1; continue; }
        int ox = bwd ? vowr_order(i - s.m, s.d.nx, R) : 4;
        const T *cx = s.ctab[ox];
        size_t id = s.idx(i, j, k);
        T dtr = dt / s.rho[id];
        T a = dtr / s.dx, b = dtr / s.dy, c = dtr / s.dz;
        T sxx_x = Dp(sxx, id, SX, cx, ox), syy_y = Dp(syy, id, SY, cy, oy), szz_z = Dp(szz, id, SZ, cz, oz);
        T syx_x = Dm(sxy, id, SX, cx, ox), sxy_y = Dm(sxy, id, SY, cy, oy);
        T sxz_z = Dm(sxz, id, SZ, cz, oz), szx_x = Dm(sxz, id, SX, cx, ox);
        T szy_y = Dm(syz, id, SY, cy, oy), syz_z = Dm(syz, id, SZ, cz, oz);
        vx[id] = s.qv(vx[id] + (a * sxx_x + b * sxy_y + c * sxz_z));
        vy[id] = s.qv(vy[id] + (a * syx_x + b * syy_y + c * syz_z));
        vz[id] = s.qv(vz[id] + (a * szx_x + b * szy_y + c * szz_z));
        if (rtm_p) {
          T sx = Dp(sau, id, SX, cx, ox), sy = Dp(sau, id, SY, cy, oy), sz = Dp(sau, id, SZ, cz, oz);
          vxp[id] = s.qv(vxp[id] + a * sx);
          vyp[id] = s.qv(vyp[id] + b * sy);
          vzp[id] = s.qv(vzp[id] + c * sz);
        } else if (fwi) {
          // PHY/fdtd3d_phy_mpi_2_vel_kernel.hpp:70-78: displacement and dt times
          // the velocity increment
          ux[id] += dt * vx[id];
          uy[id] += dt * vy[id];
          uz[id] += dt * vz[id];
          vxrt[id] = dt * (a * sxx_x + b * sxy_y + c * sxz_z);
          vyrt[id] = dt * (a * syx_x + b * syy_y + c * syz_z);
          vzrt[id] = dt * (a * szx_x + b * szy_y + c * szz_z);
        }
      }
    }
}

// PHY/fdtd3d_phy_mpi_3_stress_kernel.hpp:7-60 (inside) and edges :62-1306.
template <class T>
static void stress_phy(Sim<T> &s, Wave<T> &w, int mode, T dt, int R) {
  const bool bwd = mode == RTM_BACKWARD;
  // sau in the RTM modes only (PHY/fdtd3d_phy_mpi_3_stress_kernel.hpp:53-58)
  const bool with_sau = mode == RTM_FORWARD || mode == RTM_BACKWARD || mode == RTM_REVERSE;
  const T *vx = w.f[VX].data(), *vy = w.f[VY].data(), *vz = w.f[VZ].data();
  T *sxx = w.f[SXX].data(), *syy = w.f[SYY].data(), *szz = w.f[SZZ].data();
  T *sxy = w.f[SXY].data(), *sxz = w.f[SXZ].data(), *syz = w.f[SYZ].data();
  T *sau = w.f[SAU].data();
  const size_t SX = 1, SY = s.NX, SZ = s.NXY;
  const T a = dt / s.dx, b = dt / s.dy, c = dt / s.dz;
  const bool fast = s.fmt == STORE_F32 && g_fault == F_NONE && s.d.nx >= 8 && !g_scalar;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = s.m; k < s.m + s.d.nz; k++)
    for (int j = s.m; j < s.m + s.d.ny; j++) {
      int oz = bwd ? vowr_order(k - s.m, s.d.nz, R) : 4;
      int oy = bwd ? vowr_order(j - s.m, s.d.ny, R) : 4;
      const T *cz = s.ctab[oz], *cy = s.ctab[oy];
      // [f0, f1): the row's full-order run, updated by stress_run4
      int f0 = s.m + s.d.nx, f1 = f0;
      if (fast && oy == 4 && oz == 4) {
        f0 = s.m + (bwd ? 4 : 0);
        f1 = s.m + s.d.nx - (bwd ? 4 : 0);
        if (with_sau) stress_run4<T, true>(s, w, dt, s.idx(f0, j, k), f1 - f0);
        else stress_run4<T, false>(s, w, dt, s.idx(f0, j, k), f1 - f0);
      }
      for (int i = s.m; i < s.m + s.d.nx; i++) {
        if (i == f0) { i = f1 - 1; continue; }
        int ox = bwd ? vowr_order(i - s.m, s.d.nx, R) : 4;
        const T *cx = s.ctab[ox];
        size_t id = s.idx(i, j, k);
        T lm = s.lam[id], mu = s.mu[id];
        T l2m = lm + T(2) * mu;
        T vx_x = Dm(vx, id, SX, cx, ox), vx_y = Dp(vx, id, SY, cy, oy), vx_z = Dp(vx, id, SZ, cz, oz);
        T vy_x = Dp(vy, id, SX, cx, ox), vy_y = Dm(vy, id, SY, cy, oy), vy_z = Dp(vy, id, SZ, cz, oz);
        T vz_x = Dp(vz, id, SX, cx, ox), vz_y = Dp(vz, id, SY, cy, oy), vz_z = Dm(vz, id, SZ, cz, oz);
        sxx[id] = s.qs(sxx[id] + (a * l2m * vx_x + b * lm * vy_y + c * lm * vz_z));
        syy[id] = s.qs(syy[id] + (a * lm * vx_x + b * l2m * vy_y + c * lm * vz_z));
        szz[id] = s.qs(szz[id] + (a * lm * vx_x + b * lm * vy_y + c * l2m * vz_z));
        const T mxy = g_fault == F_SHEAR_LAMBDA ? lm : mu;
        sxy[id] = s.qs(sxy[id] + (a * mxy * vy_x + b * mxy * vx_y));
        sxz[id] = s.qs(sxz[id] + (a * mu * vz_x + c * mu * vx_z));
        syz[id] = s.qs(syz[id] + (b * mu * vz_y + c * mu * vy_z));
        if (with_sau) sau[id] = s.qs(sau[id] + (a * l2m * vx_x + b * l2m * vy_y + c * l2m * vz_z));
      }
    }
}

// CPML runs of one x row segment whose damped axes (X, Y, Z) are the same
// for every cell, written so the compiler can vectorize them across i: per
// cell, exactly the operations of the per-cell code in vel_pml / stress_pml
// below, in the same order.  Used for storage without quantization and no
// planted mistake (g_scalar forces the per-cell code everywhere).

template <class T>
__attribute__((always_inline)) static inline T Dp2(const T *__restrict f, size_t i, size_t s, T c0, T c1) {
  T acc = 0;
  acc += c0 * (f[i + s] - f[i]);
  acc += c1 * (f[i + 2 * s] - f[i - s]);
  return acc;
}
template <class T>
__attribute__((always_inline)) static inline T Dm2(const T *__restrict f, size_t i, size_t s, T c0, T c1) {
  T acc = 0;
  acc += c0 * (f[i] - f[i - s]);
  acc += c1 * (f[i + s] - f[i - 2 * s]);
  return acc;
}

template <class T, bool X, bool Y, bool Z>
static void vel_pml_run(Sim<T> &s, Wave<T> &w, T dt, size_t id0, int n) {
  T *__restrict vx = w.f[VX].data(), *__restrict vy = w.f[VY].data(), *__restrict vz = w.f[VZ].data();
  const T *__restrict sxx = w.f[SXX].data(), *__restrict syy = w.f[SYY].data(), *__restrict szz = w.f[SZZ].data();
  const T *__restrict sxy = w.f[SXY].data(), *__restrict sxz = w.f[SXZ].data(), *__restrict syz = w.f[SYZ].data();
  T *__restrict p0 = w.p[P_SXX_X].data(), *__restrict p1 = w.p[P_SXY_X].data(), *__restrict p2 = w.p[P_SXZ_X].data();
  T *__restrict p3 = w.p[P_SYY_Y].data(), *__restrict p4 = w.p[P_SXY_Y].data(), *__restrict p5 = w.p[P_SYZ_Y].data();
  T *__restrict p6 = w.p[P_SZZ_Z].data(), *__restrict p7 = w.p[P_SXZ_Z].data(), *__restrict p8 = w.p[P_SYZ_Z].data();
  const T *__restrict ax = s.ax.data(), *__restrict bx = s.bx.data(), *__restrict axh = s.axh.data(),
          *__restrict bxh = s.bxh.data();
  const T *__restrict ay = s.ay.data(), *__restrict by = s.by.data(), *__restrict ayh = s.ayh.data(),
          *__restrict byh = s.byh.data();
  const T *__restrict az = s.az.data(), *__restrict bz = s.bz.data(), *__restrict azh = s.azh.data(),
          *__restrict bzh = s.bzh.data();
  const T *__restrict rho = s.rho.data();
  const size_t SX = 1, SY = s.NX, SZ = s.NXY;
  const T c0 = s.c4[0], c1 = s.c4[1], hx = s.dx, hy = s.dy, hz = s.dz;
#pragma omp simd
  for (int q = 0; q < n; q++) {
    const size_t id = id0 + (size_t)q;
    T dtr = dt / rho[id];
    T sxx_x = Dp2(sxx, id, SX, c0, c1), syx_x = Dm2(sxy, id, SX, c0, c1), szx_x = Dm2(sxz, id, SX, c0, c1);
    T syy_y = Dp2(syy, id, SY, c0, c1), sxy_y = Dm2(sxy, id, SY, c0, c1), szy_y = Dm2(syz, id, SY, c0, c1);
    T szz_z = Dp2(szz, id, SZ, c0, c1), sxz_z = Dm2(sxz, id, SZ, c0, c1), syz_z = Dm2(syz, id, SZ, c0, c1);
    sxx_x = sxx_x / hx; syy_y = syy_y / hy; szz_z = szz_z / hz;
    sxy_y = sxy_y / hy; sxz_z = sxz_z / hz; syx_x = syx_x / hx;
    syz_z = syz_z / hz; szx_x = szx_x / hx; szy_y = szy_y / hy;
    if (X) {
      T q0 = bxh[id] * p0[id] + axh[id] * sxx_x, q1 = bx[id] * p1[id] + ax[id] * syx_x;
      T q2 = bx[id] * p2[id] + ax[id] * szx_x;
      p0[id] = q0; p1[id] = q1; p2[id] = q2;
      sxx_x += q0; syx_x += q1; szx_x += q2;
    }
    if (Y) {
      T q3 = byh[id] * p3[id] + ayh[id] * syy_y, q4 = by[id] * p4[id] + ay[id] * sxy_y;
      T q5 = by[id] * p5[id] + ay[id] * szy_y;
      p3[id] = q3; p4[id] = q4; p5[id] = q5;
      syy_y += q3; sxy_y += q4; szy_y += q5;
    }
    if (Z) {
      T q6 = bzh[id] * p6[id] + azh[id] * szz_z, q7 = bz[id] * p7[id] + az[id] * sxz_z;
      T q8 = bz[id] * p8[id] + az[id] * syz_z;
      p6[id] = q6; p7[id] = q7; p8[id] = q8;
      szz_z += q6; sxz_z += q7; syz_z += q8;
    }
    vx[id] = vx[id] + dtr * (sxx_x + sxy_y + sxz_z);
    vy[id] = vy[id] + dtr * (syx_x + syy_y + syz_z);
    vz[id] = vz[id] + dtr * (szx_x + szy_y + szz_z);
  }
}

template <class T, bool X, bool Y, bool Z>
static void stress_pml_run(Sim<T> &s, Wave<T> &w, T dt, size_t id0, int n) {
  const T *__restrict vx = w.f[VX].data(), *__restrict vy = w.f[VY].data(), *__restrict vz = w.f[VZ].data();
  T *__restrict sxx = w.f[SXX].data(), *__restrict syy = w.f[SYY].data(), *__restrict szz = w.f[SZZ].data();
  T *__restrict sxy = w.f[SXY].data(), *__restrict sxz = w.f[SXZ].data(), *__restrict syz = w.f[SYZ].data();
  T *__restrict sau = w.f[SAU].data();
  T *__restrict p0 = w.p[P_VX_X].data(), *__restrict p1 = w.p[P_VY_X].data(), *__restrict p2 = w.p[P_VZ_X].data();
  T *__restrict p3 = w.p[P_VY_Y].data(), *__restrict p4 = w.p[P_VX_Y].data(), *__restrict p5 = w.p[P_VZ_Y].data();
  T *__restrict p6 = w.p[P_VZ_Z].data(), *__restrict p7 = w.p[P_VX_Z].data(), *__restrict p8 = w.p[P_VY_Z].data();
  const T *__restrict ax = s.ax.data(), *__restrict bx = s.bx.data(), *__restrict axh = s.axh.data(),
          *__restrict bxh = s.bxh.data();
  const T *__restrict ay = s.ay.data(), *__restrict by = s.by.data(), *__restrict ayh = s.ayh.data(),
          *__restrict byh = s.byh.data();
  const T *__restrict az = s.az.data(), *__restrict bz = s.bz.data(), *__restrict azh = s.azh.data(),
          *__restrict bzh = s.bzh.data();
  const T *__restrict lam = s.lam.data(), *__restrict muv = s.mu.data();
  const size_t SX = 1, SY = s.NX, SZ = s.NXY;
  const T c0 = s.c4[0], c1 = s.c4[1], hx = s.dx, hy = s.dy, hz = s.dz;
#pragma omp simd
  for (int q = 0; q < n; q++) {
    const size_t id = id0 + (size_t)q;
    T vx_x = Dm2(vx, id, SX, c0, c1), vy_x = Dp2(vy, id, SX, c0, c1), vz_x = Dp2(vz, id, SX, c0, c1);
    T vy_y = Dm2(vy, id, SY, c0, c1), vx_y = Dp2(vx, id, SY, c0, c1), vz_y = Dp2(vz, id, SY, c0, c1);
    T vz_z = Dm2(vz, id, SZ, c0, c1), vx_z = Dp2(vx, id, SZ, c0, c1), vy_z = Dp2(vy, id, SZ, c0, c1);
    T lm = lam[id], mu = muv[id];
    T l2m = lm + T(2) * mu;
    vx_x = vx_x / hx; vy_y = vy_y / hy; vz_z = vz_z / hz;
    vy_x = vy_x / hx; vz_x = vz_x / hx; vx_y = vx_y / hy;
    vz_y = vz_y / hy; vx_z = vx_z / hz; vy_z = vy_z / hz;
    if (X) {
      T q0 = bx[id] * p0[id] + ax[id] * vx_x, q1 = bxh[id] * p1[id] + axh[id] * vy_x;
      T q2 = bxh[id] * p2[id] + axh[id] * vz_x;
      p0[id] = q0; p1[id] = q1; p2[id] = q2;
      vx_x += q0; vy_x += q1; vz_x += q2;
    }
    if (Y) {
      T q3 = by[id] * p3[id] + ay[id] * vy_y, q4 = byh[id] * p4[id] + ayh[id] * vx_y;
      T q5 = byh[id] * p5[id] + ayh[id] * vz_y;
      p3[id] = q3; p4[id] = q4; p5[id] = q5;
      vy_y += q3; vx_y += q4; vz_y += q5;
    }
    if (Z) {
      T q6 = bz[id] * p6[id] + az[id] * vz_z, q7 = bzh[id] * p7[id] + azh[id] * vx_z;
      T q8 = bzh[id] * p8[id] + azh[id] * vy_z;
      p6[id] = q6; p7[id] = q7; p8[id] = q8;
      vz_z += q6; vx_z += q7; vy_z += q8;
    }
    sxx[id] = sxx[id] + dt * (l2m * vx_x + lm * vy_y + lm * vz_z);
    syy[id] = syy[id] + dt * (lm * vx_x + l2m * vy_y + lm * vz_z);
    szz[id] = szz[id] + dt * (lm * vx_x + lm * vy_y + l2m * vz_z);
    sau[id] = sau[id] + dt * (l2m * vx_x + l2m * vy_y + l2m * vz_z);
    sxy[id] = sxy[id] + dt * mu * (vy_x + vx_y);
    sxz[id] = sxz[id] + dt * mu * (vz_x + vx_z);
    syz[id] = syz[id] + dt * mu * (vz_y + vy_z);
  }
}

// Run one segment with damped-axis flags (x, y, z) known at run time.
template <class T, bool VEL>
static void pml_run(Sim<T> &s, Wave<T> &w, T dt, size_t id0, int n, bool x, bool y, bool z) {
  if (n <= 0) return;
#define DDX_ORC_RUN(X, Y, Z)                                       \
  if (x == X && y == Y && z == Z) {                                \
    if (VEL) vel_pml_run<T, X, Y, Z>(s, w, dt, id0, n);            \
    else stress_pml_run<T, X, Y, Z>(s, w, dt, id0, n);             \
    return;                                                        \
  }
  DDX_ORC_RUN(false, false, false) DDX_ORC_RUN(false, false, true) DDX_ORC_RUN(false, true, false)
  DDX_ORC_RUN(false, true, true) DDX_ORC_RUN(true, false, false) DDX_ORC_RUN(true, false, true)
  DDX_ORC_RUN(true, true, false) DDX_ORC_RUN(true, true, true)
#undef DDX_ORC_RUN
}

// The CPML cells of row (j, k) as runs of pml_run: [2, m) and [N - m, N - 2)
// are damped along x, [m, N - m) is not and belongs to the physical region
// when the row does.
template <class T, bool VEL>
static void pml_row(Sim<T> &s, Wave<T> &w, T dt, int j, int k) {
  const bool yin = Sim<T>::in_pml_axis(j, s.NY, s.m), zin = Sim<T>::in_pml_axis(k, s.NZ, s.m);
  const bool row_in = j >= s.m && j < s.m + s.d.ny && k >= s.m && k < s.m + s.d.nz;
  const int m = s.m, e = s.m + s.d.nx;
  pml_run<T, VEL>(s, w, dt, s.idx(2, j, k), m - 2, true, yin, zin);
  if (!row_in) pml_run<T, VEL>(s, w, dt, s.idx(m, j, k), e - m, false, yin, zin);
  pml_run<T, VEL>(s, w, dt, s.idx(e, j, k), s.NX - 2 - e, true, yin, zin);
}

// CPML/fdtd3d_cpml_mpi_2_kernel_device_{x,y,z}_meat.hpp, velocity halves
// (z: :9-195).  4th order, derivative divided by the spacing, then the
// memory variable psi = b psi + a d, d += psi, then v += dt/rho * sum.
// A slab damps an axis exactly where the cell lies in that axis' damping
// range (other cells of the slab carry a = 0 and keep psi = 0).
template <class T>
static void vel_pml(Sim<T> &s, Wave<T> &w, T dt) {
  T *vx = w.f[VX].data(), *vy = w.f[VY].data(), *vz = w.f[VZ].data();
  const T *sxx = w.f[SXX].data(), *syy = w.f[SYY].data(), *szz = w.f[SZZ].data();
  const T *sxy = w.f[SXY].data(), *sxz = w.f[SXZ].data(), *syz = w.f[SYZ].data();
  T **p = nullptr;
  T *pp[NPSI];
  for (int q = 0; q < NPSI; q++) pp[q] = w.p[q].data();
  p = pp;
  const size_t SX = 1, SY = s.NX, SZ = s.NXY;
  const T *c = s.c4;
  const bool fast = s.fmt == STORE_F32 && g_fault == F_NONE && !g_scalar;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 2; k < s.NZ - 2; k++)
    for (int j = 2; j < s.NY - 2; j++) {
      if (fast) { pml_row<T, true>(s, w, dt, j, k); continue; }
      const bool row_in = j >= s.m && j < s.m + s.d.ny && k >= s.m && k < s.m + s.d.nz;
      for (int i = 2; i < s.NX - 2; i++) {
        if (row_in && i == s.m) { i = s.m + s.d.nx - 1; continue; }  // skip the physical region
        bool xin = Sim<T>::in_pml_axis(i, s.NX, s.m);
        bool yin = Sim<T>::in_pml_axis(j, s.NY, s.m);
        bool zin = Sim<T>::in_pml_axis(k, s.NZ, s.m);
        size_t id = s.idx(i, j, k);
        T dtr = dt / s.rho[id];
        T sxx_x = Dp(sxx, id, SX, c, 2), syx_x = Dm(sxy, id, SX, c, 2), szx_x = Dm(sxz, id, SX, c, 2);
        T syy_y = Dp(syy, id, SY, c, 2), sxy_y = Dm(sxy, id, SY, c, 2), szy_y = Dm(syz, id, SY, c, 2);
        T szz_z = Dp(szz, id, SZ, c, 2), sxz_z = Dm(sxz, id, SZ, c, 2), syz_z = Dm(syz, id, SZ, c, 2);
        sxx_x = sxx_x / s.dx; syy_y = syy_y / s.dy; szz_z = szz_z / s.dz;
        sxy_y = sxy_y / s.dy; sxz_z = sxz_z / s.dz; syx_x = syx_x / s.dx;
        syz_z = syz_z / s.dz; szx_x = szx_x / s.dx; szy_y = szy_y / s.dy;
        if (xin) {
          if (g_fault == F_PML_VX_X)
            p[P_SXX_X][id] = s.qs(s.bx[id] * p[P_SXX_X][id] + s.ax[id] * sxx_x);
          else
            p[P_SXX_X][id] = s.qs(s.bxh[id] * p[P_SXX_X][id] + s.axh[id] * sxx_x);
          if (g_fault == F_PML_VY_SXY_X)
            p[P_SXY_X][id] = s.qs(s.bxh[id] * p[P_SXY_X][id] + s.axh[id] * syx_x);
          else
            p[P_SXY_X][id] = s.qs(s.bx[id] * p[P_SXY_X][id] + s.ax[id] * syx_x);
          p[P_SXZ_X][id] = s.qs(s.bx[id] * p[P_SXZ_X][id] + s.ax[id] * szx_x);
          sxx_x += p[P_SXX_X][id]; syx_x += p[P_SXY_X][id]; szx_x += p[P_SXZ_X][id];
        }
        if (yin) {
          if (g_fault == F_PML_VY_Y)
            p[P_SYY_Y][id] = s.qs(s.by[id] * p[P_SYY_Y][id] + s.ay[id] * syy_y);
          else
            p[P_SYY_Y][id] = s.qs(s.byh[id] * p[P_SYY_Y][id] + s.ayh[id] * syy_y);
          p[P_SXY_Y][id] = s.qs(s.by[id] * p[P_SXY_Y][id] + s.ay[id] * sxy_y);
          p[P_SYZ_Y][id] = s.qs(s.by[id] * p[P_SYZ_Y][id] + s.ay[id] * szy_y);
          syy_y += p[P_SYY_Y][id]; sxy_y += p[P_SXY_Y][id]; szy_y += p[P_SYZ_Y][id];
        }
        if (zin) {
          if (g_fault == F_PML_VZ_Z)
            p[P_SZZ_Z][id] = s.qs(s.bz[id] * p[P_SZZ_Z][id] + s.az[id] * szz_z);
          else
            p[P_SZZ_Z][id] = s.qs(s.bzh[id] * p[P_SZZ_Z][id] + s.azh[id] * szz_z);
          p[P_SXZ_Z][id] = s.qs(s.bz[id] * p[P_SXZ_Z][id] + s.az[id] * sxz_z);
          p[P_SYZ_Z][id] = s.qs(s.bz[id] * p[P_SYZ_Z][id] + s.az[id] * syz_z);
          szz_z += p[P_SZZ_Z][id]; sxz_z += p[P_SXZ_Z][id]; syz_z += p[P_SYZ_Z][id];
        }
        vx[id] = s.qv(vx[id] + dtr * (sxx_x + sxy_y + sxz_z));
        vy[id] = s.qv(vy[id] + dtr * (syx_x + syy_y + syz_z));
        vz[id] = s.qv(vz[id] + dtr * (szx_x + szy_y + szz_z));
      }
    }
}

// CPML/..._device_z_meat.hpp:197-391 (stress half).  sau is updated in the
// CPML in every mode (:387).
template <class T>
static void stress_pml(Sim<T> &s, Wave<T> &w, T dt) {
  const T *vx = w.f[VX].data(), *vy = w.f[VY].data(), *vz = w.f[VZ].data();
  T *sxx = w.f[SXX].data(), *syy = w.f[SYY].data(), *szz = w.f[SZZ].data();
  T *sxy = w.f[SXY].data(), *sxz = w.f[SXZ].data(), *syz = w.f[SYZ].data();
  T *sau = w.f[SAU].data();
  T *p[NPSI];
  for (int q = 0; q < NPSI; q++) p[q] = w.p[q].data();
  const size_t SX = 1, SY = s.NX, SZ = s.NXY;
  const T *c = s.c4;
  const bool fast = s.fmt == STORE_F32 && g_fault == F_NONE && !g_scalar;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 2; k < s.NZ - 2; k++)
    for (int j = 2; j < s.NY - 2; j++) {
      if (fast) { pml_row<T, false>(s, w, dt, j, k); continue; }
      const bool row_in = j >= s.m && j < s.m + s.d.ny && k >= s.m && k < s.m + s.d.nz;
      for (int i = 2; i < s.NX - 2; i++) {
        if (row_in && i == s.m) { i = s.m + s.d.nx - 1; continue; }  // skip the physical region
        bool xin = Sim<T>::in_pml_axis(i, s.NX, s.m);
        bool yin = Sim<T>::in_pml_axis(j, s.NY, s.m);
        bool zin = Sim<T>::in_pml_axis(k, s.NZ, s.m);
        size_t id = s.idx(i, j, k);
        T vx_x = Dm(vx, id, SX, c, 2), vy_x = Dp(vy, id, SX, c, 2), vz_x = Dp(vz, id, SX, c, 2);
        T vy_y = Dm(vy, id, SY, c, 2), vx_y = Dp(vx, id, SY, c, 2), vz_y = Dp(vz, id, SY, c, 2);
        T vz_z = Dm(vz, id, SZ, c, 2), vx_z = Dp(vx, id, SZ, c, 2), vy_z = Dp(vy, id, SZ, c, 2);
        T lm = s.lam[id], mu = s.mu[id];
        T l2m = lm + T(2) * mu;
        vx_x = vx_x / s.dx; vy_y = vy_y / s.dy; vz_z = vz_z / s.dz;
        vy_x = vy_x / s.dx; vz_x = vz_x / s.dx; vx_y = vx_y / s.dy;
        vz_y = vz_y / s.dy; vx_z = vx_z / s.dz; vy_z = vy_z / s.dz;
        if (xin) {
          if (g_fault == F_PML_SXX_X)
            p[P_VX_X][id] = s.qv(s.bxh[id] * p[P_VX_X][id] + s.axh[id] * vx_x);
          else
            p[P_VX_X][id] = s.qv(s.bx[id] * p[P_VX_X][id] + s.ax[id] * vx_x);
          if (g_fault == F_PML_SXY_X)
            p[P_VY_X][id] = s.qv(s.bx[id] * p[P_VY_X][id] + s.ax[id] * vy_x);
          else
            p[P_VY_X][id] = s.qv(s.bxh[id] * p[P_VY_X][id] + s.axh[id] * vy_x);
          p[P_VZ_X][id] = s.qv(s.bxh[id] * p[P_VZ_X][id] + s.axh[id] * vz_x);
          vx_x += p[P_VX_X][id]; vy_x += p[P_VY_X][id]; vz_x += p[P_VZ_X][id];
        }
        if (yin) {
          p[P_VY_Y][id] = s.qv(s.by[id] * p[P_VY_Y][id] + s.ay[id] * vy_y);
          p[P_VX_Y][id] = s.qv(s.byh[id] * p[P_VX_Y][id] + s.ayh[id] * vx_y);
          p[P_VZ_Y][id] = s.qv(s.byh[id] * p[P_VZ_Y][id] + s.ayh[id] * vz_y);
          vy_y += p[P_VY_Y][id]; vx_y += p[P_VX_Y][id]; vz_y += p[P_VZ_Y][id];
        }
        if (zin) {
          p[P_VZ_Z][id] = s.qv(s.bz[id] * p[P_VZ_Z][id] + s.az[id] * vz_z);
          p[P_VX_Z][id] = s.qv(s.bzh[id] * p[P_VX_Z][id] + s.azh[id] * vx_z);
          p[P_VY_Z][id] = s.qv(s.bzh[id] * p[P_VY_Z][id] + s.azh[id] * vy_z);
          vz_z += p[P_VZ_Z][id]; vx_z += p[P_VX_Z][id]; vy_z += p[P_VY_Z][id];
        }
        sxx[id] = s.qs(sxx[id] + dt * (l2m * vx_x + lm * vy_y + lm * vz_z));
        syy[id] = s.qs(syy[id] + dt * (lm * vx_x + l2m * vy_y + lm * vz_z));
        szz[id] = s.qs(szz[id] + dt * (lm * vx_x + lm * vy_y + l2m * vz_z));
        sau[id] = s.qs(sau[id] + dt * (l2m * vx_x + l2m * vy_y + l2m * vz_z));
        sxy[id] = s.qs(sxy[id] + dt * mu * (vy_x + vx_y));
        sxz[id] = s.qs(sxz[id] + dt * mu * (vz_x + vx_z));
        syz[id] = s.qs(syz[id] + dt * mu * (vz_y + vy_z));
      }
    }
}

// One velocity half step of a set (phy + CPML; backward mode is phy only,
// PHY/fdtd3d_phy_mpi_4_device_meat.hpp:227-256, BLEND/...meat.hpp:161-170).
template <class T>
static void vel_step(Sim<T> &s, Wave<T> &w, int mode) {
  T dt = mode == RTM_BACKWARD ? -s.dt : s.dt;
  vel_phy(s, w, mode, dt, s.d.rec_width);
  if (mode != RTM_BACKWARD) vel_pml(s, w, dt);
}
template <class T>
static void stress_step(Sim<T> &s, Wave<T> &w, int mode) {
  T dt = mode == RTM_BACKWARD ? -s.dt : s.dt;
  stress_phy(s, w, mode, dt, s.d.rec_width);
  if (mode != RTM_BACKWARD) stress_pml(s, w, dt);
}

// PHY/fdtd3d_phy_mpi_3_stress_kernel.hpp:1308-1364: explosive source into the
// three normal stresses (sign +1 add, -1 subtract).  src_mode 1 injects szz
// only (the "vertical-stress component" variant).
template <class T>
static void add_source(Sim<T> &s, Wave<T> &w, long node, T val, int sign, int src_mode) {
  if (node < 0) return;
  T *sxx = w.f[SXX].data(), *syy = w.f[SYY].data(), *szz = w.f[SZZ].data();
  if (sign > 0) {
    if (src_mode == 0) { sxx[node] = s.qs(sxx[node] + val); syy[node] = s.qs(syy[node] + val); }
    if (src_mode == 1 && g_fault == F_SRC_SZZ) sxx[node] = s.qs(sxx[node] + val);
    else szz[node] = s.qs(szz[node] + val);
  } else {
    if (src_mode == 0) { sxx[node] = s.qs(sxx[node] - val); syy[node] = s.qs(syy[node] - val); }
    szz[node] = s.qs(szz[node] - val);
  }
}

// Ring boxes (BASE/fdtd3d_base_5_wave_recon_meat.hpp:14-44 and
// OP/jarvis_1_mpicu_halo_store_operator_meat.hpp:122-135): faces of the
// global gridphy, just outside it.  Layout of one time slice (the oracle's
// own): for field in (vx,vy,vz,sxx,syy,szz,sau,sxy,sxz,syz), side 0..5,
// box x-fastest.
struct Box { int i0, j0, k0, ni, nj, nk; };
template <class T>
static Box ring_box(const Sim<T> &s, int side, int wdt) {
  int m = s.m, nx = s.d.nx, ny = s.d.ny, nz = s.d.nz;
  switch (side) {
    case 0: return {m, m, m - wdt, nx, ny, wdt};
    case 1: return {m, m, m + nz, nx, ny, wdt};
    case 2: return {m, m - wdt, m, nx, wdt, nz};
    case 3: return {m, m + ny, m, nx, wdt, nz};
    case 4: return {m - wdt, m, m, wdt, ny, nz};
    default: return {m + nx, m, m, wdt, ny, nz};
  }
}
static const int kRingVel[3] = {VX, VY, VZ};
static const int kRingStress[7] = {SXX, SYY, SZZ, SAU, SXY, SXZ, SYZ};

template <class T>
static size_t ring_slice_size(const Sim<T> &s) {
  size_t n = 0;
  for (int f : {VX, VY, VZ, SXX, SYY, SZZ, SAU, SXY, SXZ, SYZ})
    for (int side = 0; side < 6; side++) {
      Box b = ring_box(s, side, ring_width(f, side, s.d.rec_width));
      n += (size_t)b.ni * b.nj * b.nk;
    }
  return n;
}
// Scale exponent of one field's slice under per-slice scaling: its largest
// |value| m lands in [2^14, 2^15), below the binary16 maximum.
static int ring_exponent(float m) {
  if (!(m > 0.f)) return 0;
  int ex = 0;
  std::frexp(m, &ex);
  return std::min(100, std::max(-100, 15 - ex));
}

// STORE_BFP16 save of one ring box into the oracle's slice layout (box
// x-fastest from off): the box splits into the subdomains' parts along its
// tangential axes (those spanning the whole gridphy), and each part's cells,
// x-fastest from its own origin, form runs of kBfpRun.
template <class T>
static void ring_save_bfp(const Sim<T> &s, const T *fld, const Box &b, T *slice, size_t off) {
  const int lo[3] = {b.i0, b.j0, b.k0}, n[3] = {b.ni, b.nj, b.nk};
  const int pn[3] = {s.d.nx, s.d.ny, s.d.nz};
  std::vector<int> edge[3];
  for (int a = 0; a < 3; a++) {
    edge[a].push_back(lo[a]);
    if (lo[a] == s.m && n[a] == pn[a])
      for (int c : s.rcut[a])
        if (c > 0 && c < pn[a]) edge[a].push_back(s.m + c);
    edge[a].push_back(lo[a] + n[a]);
  }
  for (size_t pz = 0; pz + 1 < edge[2].size(); pz++)
    for (size_t py = 0; py + 1 < edge[1].size(); py++)
      for (size_t px = 0; px + 1 < edge[0].size(); px++) {
        const int x0 = edge[0][px], y0 = edge[1][py], z0 = edge[2][pz];
        const int nx = edge[0][px + 1] - x0, ny = edge[1][py + 1] - y0, nz = edge[2][pz + 1] - z0;
        const long cnt = (long)nx * ny * nz;
        auto cell = [&](long e, size_t &id, size_t &q) {
          const int i = (int)(e % nx), j = (int)(e / nx % ny), k = (int)(e / ((long)nx * ny));
          id = s.idx(x0 + i, y0 + j, z0 + k);
          q = off + (size_t)(x0 + i - b.i0) + (size_t)b.ni * ((size_t)(y0 + j - b.j0) + (size_t)b.nj * (z0 + k - b.k0));
        };
        for (long r0 = 0; r0 < cnt; r0 += kBfpRun) {
          const long r1 = std::min(cnt, r0 + kBfpRun);
          float mx = 0.f;
          size_t id, q;
          for (long e = r0; e < r1; e++) {
            cell(e, id, q);
            mx = std::max(mx, (float)std::fabs((float)fld[id]));
          }
          const int sx = bfp_exponent(mx);
          for (long e = r0; e < r1; e++) {
            cell(e, id, q);
            slice[q] = (T)bfp_round((float)fld[id], sx);
          }
        }
      }
}

// dir = +1 gather into slice, -1 scatter from slice; which: 0 all, 1 vel, 2 stress
template <class T>
static void ring_copy(Sim<T> &s, Wave<T> &w, T *slice, int dir, int which) {
  size_t off = 0;
  for (int f : {VX, VY, VZ, SXX, SYY, SZZ, SAU, SXY, SXZ, SYZ}) {
    bool isvel = f <= VZ;
    float fsc = isvel ? s.rsv : s.rss;
    if (dir > 0 && s.ring_fmt == STORE_BFP16) {
      for (int side = 0; side < 6; side++) {
        Box b = ring_box(s, side, ring_width(f, side, s.d.rec_width));
        ring_save_bfp(s, w.f[f].data(), b, slice, off);
        off += (size_t)b.ni * b.nj * b.nk;
      }
      continue;
    }
    if (dir > 0 && s.ring_dyn) {  // the field's largest |value| over its boxes of this slice
      float m = 0.f;
      for (int side = 0; side < 6; side++) {
        Box b = ring_box(s, side, ring_width(f, side, s.d.rec_width));
        for (int k = 0; k < b.nk; k++)
          for (int j = 0; j < b.nj; j++)
            for (int i = 0; i < b.ni; i++)
              m = std::max(m, (float)std::fabs(w.f[f][s.idx(b.i0 + i, b.j0 + j, b.k0 + k)]));
      }
      fsc = std::ldexp(1.0f, ring_exponent(m));
    }
    for (int side = 0; side < 6; side++) {
      Box b = ring_box(s, side, ring_width(f, side, s.d.rec_width));
      size_t cnt = (size_t)b.ni * b.nj * b.nk;
      bool act = which == 0 || (which == 1 && isvel) || (which == 2 && !isvel);
      if (act) {
        size_t q = off;
        for (int k = 0; k < b.nk; k++)
          for (int j = 0; j < b.nj; j++)
            for (int i = 0; i < b.ni; i++, q++) {
              size_t id = s.idx(b.i0 + i, b.j0 + j, b.k0 + k);
              if (dir > 0) slice[q] = quant(s.ring_fmt, w.f[f][id], fsc);
              else w.f[f][id] = slice[q];
            }
      }
      off += cnt;
    }
  }
}

// IMG/rtm3d_iamge_kernel.hpp:6-51 (cal_correlation_from_wave).
template <class T>
static void image_step(Sim<T> &s) {
  Wave<T> &B = s.w[0], &Rv = s.w[1];
  const int nx = s.d.nx, ny = s.d.ny, nz = s.d.nz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t id = s.idx(i + s.m, j + s.m, k + s.m);
        size_t q = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
        T pfx = B.f[VXP][id], pfy = B.f[VYP][id], pfz = B.f[VZP][id];
        T sfx = B.f[VX][id] - pfx, sfy = B.f[VY][id] - pfy, sfz = B.f[VZ][id] - pfz;
        T prx = Rv.f[VXP][id], pry = Rv.f[VYP][id], prz = Rv.f[VZP][id];
        T srx = Rv.f[VX][id] - prx, sry = Rv.f[VY][id] - pry, srz = Rv.f[VZ][id] - prz;
        s.img[0][q] += pfx * pfx + pfy * pfy + pfz * pfz;
        s.img[1][q] += pfx * prx + pfy * pry + pfz * prz;
        s.img[2][q] += pfx * srx + pfy * sry + pfz * srz;
        s.img[3][q] += sfx * prx + sfy * pry + sfz * prz;
        s.img[4][q] += sfx * srx + sfy * sry + sfz * srz;
      }
}

}  // namespace orc

// =============================================================================
// C interface (ctypes).  Precision is chosen at creation (4 = float mirror,
// 8 = double).  All field, seismogram and image outputs are returned as
// double so that both variants share one signature.
// =============================================================================
using namespace orc;

struct OrcCtx {
  int prec;
  Sim<float> *sf = nullptr;
  Sim<double> *sd = nullptr;
  std::vector<float> storef;
  std::vector<double> stored;
  // what orc_rtm_shot's snapshots hold: field snap_field of set 0 in the
  // forward pass, of set snap_set in the reverse loop
  int snap_set = 0, snap_field = VX;
};

#define DISPATCH(ctx, call) ((ctx)->prec == 4 ? call(*(ctx)->sf) : call(*(ctx)->sd))

extern "C" {

typedef struct {
  int nx, ny, nz;
  float dx, dy, dz;
  int pml;
  float dt, fm;
  int ntime;
  int rec_width;
} orc_desc;

void *orc_create(const orc_desc *dd, int prec, const float *vp, const float *vs, const float *rho) {
  Desc d;
  d.nx = dd->nx; d.ny = dd->ny; d.nz = dd->nz;
  d.dx = dd->dx; d.dy = dd->dy; d.dz = dd->dz;
  d.pml = dd->pml; d.dt = dd->dt; d.fm = dd->fm; d.ntime = dd->ntime; d.rec_width = dd->rec_width;
  if (d.pml < 4 || d.rec_width < 1 || d.rec_width > 4 || d.nx < 8 || d.ny < 8 || d.nz < 8) return nullptr;
  OrcCtx *c = new OrcCtx();
  c->prec = prec == 8 ? 8 : 4;
  if (c->prec == 4) { c->sf = new Sim<float>(); init_sim(*c->sf, d, vp, vs, rho); }
  else { c->sd = new Sim<double>(); init_sim(*c->sd, d, vp, vs, rho); }
  return c;
}

// Test hooks: plant one of the Fault mistakes (0 removes it; takes effect
// for oracles created and steps run afterwards), and choose the set and field
// of the rtm snapshots.
void orc_plant_fault(int f) { g_fault = f; }
// on = 1: every update takes the per-cell code (the vectorized runs are off)
void orc_scalar_paths(int on) { g_scalar = on != 0; }
void orc_set_snap(void *p, int set, int field) {
  OrcCtx *c = (OrcCtx *)p;
  c->snap_set = set;
  c->snap_field = field;
}

void orc_destroy(void *p) {
  OrcCtx *c = (OrcCtx *)p;
  if (!c) return;
  delete c->sf;
  delete c->sd;
  delete c;
}

// GEO/jarvis_device_geometry_bones.hpp:29-49
void orc_fd_coef(int half, double *out) { fd_coef_double(half, out); }

int orc_gap_num(float l, float r, float d) { return gap_num(l, r, d); }

// BASE/fdtd3d_base_4_seisrecord_meat.hpp:273-285 (SeisRecord::initialize):
// ntime = ceil(T/dt) in float, Ricker with n0 = floor(1.3/fm/dt + 0.5).
int orc_ntime(float T, float dt) { return (int)std::ceil(T / dt); }
void orc_ricker(float fm, float dt, int ntime, float *out) {
  float dura = 1.3f / fm;
  int n0 = (int)std::floor(dura / dt + 0.5f);
  for (int i = 0; i < ntime; i++) {
    double x = kPi * (double)fm * (double)((float)(i - n0) * dt);
    double p = std::pow(x, 2);
    out[i] = (float)(1.0e4f * (1.0f - 2.0f * p) * std::exp((-1) * p));
  }
}

// Extended-grid shape and the cell index of a point given in metres (global
// gridext has origin -pml*d, FLD get_gap_num).  Returns -1 when outside.
long orc_node(const orc_desc *d, float x, float y, float z) {
  int m = d->pml;
  int NX = d->nx + 2 * m, NY = d->ny + 2 * m, NZ = d->nz + 2 * m;
  float lx = -m * d->dx, ly = -m * d->dy, lz = -m * d->dz;
  float rx = lx + (NX - 1) * d->dx, ry = ly + (NY - 1) * d->dy, rz = lz + (NZ - 1) * d->dz;
  if (!(x >= lx && y >= ly && z >= lz && x <= rx && y <= ry && z <= rz)) return -1;
  int i = gap_num(lx, x, d->dx), j = gap_num(ly, y, d->dy), k = gap_num(lz, z, d->dz);
  return (long)i + (long)NX * ((long)j + (long)NY * k);
}

void orc_reset(void *p, int set) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec == 4) c->sf->w[set].zero(); else c->sd->w[set].zero();
}
void orc_reset_images(void *p) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec == 4) for (auto &a : c->sf->img) std::fill(a.begin(), a.end(), 0.f);
  else for (auto &a : c->sd->img) std::fill(a.begin(), a.end(), 0.0);
}

void orc_vel(void *p, int set, int mode) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec == 4) vel_step(*c->sf, c->sf->w[set], mode); else vel_step(*c->sd, c->sd->w[set], mode);
}
void orc_stress(void *p, int set, int mode) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec == 4) stress_step(*c->sf, c->sf->w[set], mode); else stress_step(*c->sd, c->sd->w[set], mode);
}
void orc_source(void *p, int set, long node, float val, int sign, int src_mode) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec == 4) add_source(*c->sf, c->sf->w[set], node, val, sign, src_mode);
  else add_source(*c->sd, c->sd->w[set], node, (double)val, sign, src_mode);
}

// FWI adjoint source (BASE/fdtd3d_base_4_seisrecord_meat.hpp:109-147,
// simulate_type fwi): v[node] += obs - syn at sample it, in float.
void orc_inject_residual(void *p, int set, const long *rec_nodes, int nrec, const float *obs, const float *syn,
                         int it) {
  OrcCtx *c = (OrcCtx *)p;
  const int ntime = c->prec == 4 ? c->sf->d.ntime : c->sd->d.ntime;
  for (int r = 0; r < nrec; r++)
    for (int q = 0; q < 3; q++) {
      const size_t k = ((size_t)q * nrec + r) * ntime + it;
      const float v = obs[k] - syn[k];
      if (c->prec == 4) c->sf->w[set].f[VX + q][rec_nodes[r]] += v;
      else c->sd->w[set].f[VX + q][rec_nodes[r]] += (double)v;
    }
}

void orc_get_field(void *p, int set, int field, double *out) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec == 4) { auto &a = c->sf->w[set].f[field]; for (size_t i = 0; i < a.size(); i++) out[i] = a[i]; }
  else { auto &a = c->sd->w[set].f[field]; for (size_t i = 0; i < a.size(); i++) out[i] = a[i]; }
}
// Stores round to the storage format, as the engine's ddx_set_field does.
void orc_set_field(void *p, int set, int field, const double *in) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec == 4) {
    Sim<float> &s = *c->sf;
    auto &a = s.w[set].f[field];
    const bool vel = field <= VZ || (field >= VXP && field <= VZP);
    const bool str = field >= SXX && field <= SAU;
    for (size_t i = 0; i < a.size(); i++) {
      const float x = (float)in[i];
      a[i] = vel ? s.qv(x) : str ? s.qs(x) : x;
    }
  }
  else { auto &a = c->sd->w[set].f[field]; for (size_t i = 0; i < a.size(); i++) a[i] = in[i]; }
}
// Storage format of the FP32 mirror's wavefields: fmt 0 FP32, 1 FP16, 2 BF16,
// with power-of-two scales 2^lg_v (velocity class) and 2^lg_s (stress class).
// The FP64 oracle always stores exactly.
void orc_set_storage(void *p, int fmt, int lg_v, int lg_s) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec != 4) return;
  c->sf->fmt = fmt;
  c->sf->sv = std::ldexp(1.0f, lg_v);
  c->sf->ss = std::ldexp(1.0f, lg_s);
}
// VOWR boundary store in format fmt (1 FP16, 2 BF16) over FP32 wavefields,
// class scales 2^lg_v (vx, vy, vz) and 2^lg_s (stresses and sau).
void orc_set_ring_storage(void *p, int fmt, int lg_v, int lg_s) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec != 4) return;
  c->sf->ring_fmt = fmt;
  c->sf->rsv = std::ldexp(1.0f, lg_v);
  c->sf->rss = std::ldexp(1.0f, lg_s);
}

// Per-slice, per-field scales of the 16-bit VOWR store (on = 1) instead of
// the class scales of orc_set_ring_storage.
void orc_set_ring_dynamic(void *p, int on) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec == 4) c->sf->ring_dyn = on != 0;
}
int orc_ring_exponent(float m) { return ring_exponent(m); }

// The block-scaled 16-bit store (STORE_BFP16) over FP32 wavefields, for the
// decomposition whose subdomains start at gridphy indices cuts[0..ncx) along
// x, then ncy along y, then ncz along z (FP32 mirror only).
void orc_set_ring_bfp(void *p, const int *cuts, int ncx, int ncy, int ncz) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec != 4) return;
  c->sf->ring_fmt = STORE_BFP16;
  const int nc[3] = {ncx, ncy, ncz};
  for (int a = 0, o = 0; a < 3; o += nc[a], a++) c->sf->rcut[a].assign(cuts + o, cuts + o + nc[a]);
}
// The store's rounding of x[0..n), runs of kBfpRun from x[0], for its checks.
void orc_bfp_round(const float *x, float *out, long n) {
  for (long r0 = 0; r0 < n; r0 += kBfpRun) {
    const long r1 = std::min(n, r0 + kBfpRun);
    float mx = 0.f;
    for (long i = r0; i < r1; i++) mx = std::max(mx, std::fabs(x[i]));
    const int sx = bfp_exponent(mx);
    for (long i = r0; i < r1; i++) out[i] = bfp_round(x[i], sx);
  }
}

// The conversions themselves, for their checks: out[i] = the value x[i]
// takes when stored in format fmt at scale 2^lg; bits[i] = the 16-bit code
// of x[i] * 2^lg (fmt 1 or 2).
void orc_quantize(int fmt, int lg, const float *x, float *out, uint16_t *bits, long n) {
  const float sc = std::ldexp(1.0f, lg);
  for (long i = 0; i < n; i++) {
    out[i] = quant(fmt, x[i], sc);
    if (bits) bits[i] = fmt == STORE_BF16 ? f32_to_bf16(x[i] * sc) : f32_to_f16(x[i] * sc);
  }
}

// Coefficient arrays, for checks: which = 0..11 (ax,bx,axh,bxh,ay,...,bzh), 12 lam, 13 mu, 14 rho
void orc_get_coef(void *p, int which, double *out) {
  OrcCtx *c = (OrcCtx *)p;
  auto fetch = [&](auto &s) {
    using T = typename std::remove_reference<decltype(s.lam[0])>::type;
    std::vector<T> *arrs[15] = {&s.ax, &s.bx, &s.axh, &s.bxh, &s.ay, &s.by, &s.ayh, &s.byh,
                                &s.az, &s.bz, &s.azh, &s.bzh, &s.lam, &s.mu, &s.rho};
    auto &a = *arrs[which];
    for (size_t i = 0; i < a.size(); i++) out[i] = (double)a[i];
  };
  if (c->prec == 4) fetch(*c->sf); else fetch(*c->sd);
}

static void sample(OrcCtx *c, int set, const long *nodes, int nrec, int it, int ntime, double *seis) {
  for (int r = 0; r < nrec; r++) {
    long n = nodes[r];
    for (int q = 0; q < 3; q++) {
      double v = c->prec == 4 ? (double)c->sf->w[set].f[VX + q][n] : c->sd->w[set].f[VX + q][n];
      seis[((size_t)q * nrec + r) * ntime + it] = v;
    }
  }
}

// Receiver sample of one step into seis[3][nrec][ntime] (SeisRecord, the
// synthetic seismograms of an FWI forward pass).
void orc_sample(void *p, int set, const long *nodes, int nrec, int it, double *seis) {
  OrcCtx *c = (OrcCtx *)p;
  sample(c, set, nodes, nrec, it, c->prec == 4 ? c->sf->d.ntime : c->sd->d.ntime, seis);
}

// BLEND/fdtd3d_blend_elastic_model_mpi_meat.hpp:44-88 (mpi_sub_forward):
// per step: velocity, source, stress, receiver sampling
// (BASE/fdtd3d_base_4_seisrecord_meat.hpp:57-77).  seis: [3][nrec][nsteps].
int orc_model_shot(void *p, long shot_node, const float *wavelet, const long *rec_nodes, int nrec,
                   int nsteps, int src_mode, double *seis) {
  OrcCtx *c = (OrcCtx *)p;
  orc_reset(p, 0);
  for (int it = 0; it < nsteps; it++) {
    orc_vel(p, 0, PURE_FORWARD);
    orc_source(p, 0, shot_node, wavelet[it], +1, src_mode);
    orc_stress(p, 0, PURE_FORWARD);
    if (seis) sample(c, 0, rec_nodes, nrec, it, nsteps, seis);
  }
  return 0;
}

size_t orc_ring_slice_size(void *p) {
  OrcCtx *c = (OrcCtx *)p;
  return c->prec == 4 ? ring_slice_size(*c->sf) : ring_slice_size(*c->sd);
}

// BLEND/...meat.hpp:125-155 (mpi_sub_forward_for_restruct) then
// RTM/rtm3d_blend_elastic_model_mpi_meat.hpp:56-113 (reverse loop) for one
// shot.  obs: [3][nrec][ntime] observed data injected at the receivers.
// When snap_every > 0, snapshots on the extended grid of the field chosen by
// orc_set_snap (default vx of set 0) are written to bwd_snap after reverse
// iterations it = ntime - q*snap_every >= 1, and of set 0 to fwd_snap after
// forward steps it = ntime-1-q*snap_every (the reconstruction check compares
// the two at the same time index; both [nsnap][N]).
int orc_rtm_shot(void *p, long shot_node, const float *wavelet, const long *rec_nodes, int nrec,
                 const float *obs, int src_mode, int snap_every, double *fwd_snap, double *bwd_snap) {
  OrcCtx *c = (OrcCtx *)p;
  int ntime = c->prec == 4 ? c->sf->d.ntime : c->sd->d.ntime;
  size_t rs = orc_ring_slice_size(p);
  size_t N = c->prec == 4 ? c->sf->N : c->sd->N;
  if (c->prec == 4) c->storef.assign(rs * (size_t)std::max(ntime - 1, 1), 0.f);
  else c->stored.assign(rs * (size_t)std::max(ntime - 1, 1), 0.0);
  orc_reset(p, 0);
  for (int it = 0; it < ntime; it++) {
    orc_vel(p, 0, RTM_FORWARD);
    orc_source(p, 0, shot_node, wavelet[it], +1, src_mode);
    orc_stress(p, 0, RTM_FORWARD);
    if (it <= ntime - 2) {
      if (c->prec == 4) ring_copy(*c->sf, c->sf->w[0], c->storef.data() + rs * it, +1, 0);
      else ring_copy(*c->sd, c->sd->w[0], c->stored.data() + rs * it, +1, 0);
    }
    if (snap_every > 0 && fwd_snap && (ntime - 1 - it) % snap_every == 0) {
      // forward state after step it == backward state after reverse iteration it+1
      int q = (ntime - 1 - it) / snap_every;
      orc_get_field(p, 0, c->snap_field, fwd_snap + (size_t)q * N);
    }
  }
  orc_reset(p, 1);
  for (int it = ntime - 1; it >= 0; it--) {
    orc_source(p, 0, shot_node, wavelet[it], -1, src_mode);
    // reverse receivers (BASE/fdtd3d_base_4_seisrecord_meat.hpp:79-107)
    const int its = g_fault == F_INJ_TIME ? it - 1 : it;
    for (int r = 0; r < nrec && its >= 0; r++) {
      long n = rec_nodes[r];
      for (int q = 0; q < 3; q++) {
        float v = obs[((size_t)q * nrec + r) * ntime + its];
        if (g_fault == F_INJ_SIGN) v = -v;
        const int qd = g_fault == F_INJ_COMP ? (q + 1) % 3 : q;
        if (c->prec == 4) {
          float &f = c->sf->w[1].f[VX + qd][n];
          f = c->sf->qv(f + v);
        } else {
          c->sd->w[1].f[VX + qd][n] += (double)v;
        }
      }
    }
    orc_stress(p, 0, RTM_BACKWARD);
    if (it - 1 >= 0) {
      if (c->prec == 4) ring_copy(*c->sf, c->sf->w[0], c->storef.data() + rs * (it - 1), -1, 2);
      else ring_copy(*c->sd, c->sd->w[0], c->stored.data() + rs * (it - 1), -1, 2);
    }
    orc_vel(p, 1, RTM_REVERSE);
    orc_vel(p, 0, RTM_BACKWARD);
    if (it - 1 >= 0) {
      if (c->prec == 4) ring_copy(*c->sf, c->sf->w[0], c->storef.data() + rs * (it - 1), -1, 1);
      else ring_copy(*c->sd, c->sd->w[0], c->stored.data() + rs * (it - 1), -1, 1);
    }
    orc_stress(p, 1, RTM_REVERSE);
    if (c->prec == 4) image_step(*c->sf); else image_step(*c->sd);
    if (snap_every > 0 && bwd_snap && it >= 1 && (ntime - it) % snap_every == 0) {
      int q = (ntime - it) / snap_every;
      orc_get_field(p, c->snap_set, c->snap_field, bwd_snap + (size_t)q * N);
    }
  }
  return 0;
}

// One imaging step from the current backward (set 0) and reverse (set 1)
// fields (IMG/rtm3d_iamge_kernel.hpp:6-51).
void orc_image_accumulate(void *p) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec == 4) image_step(*c->sf); else image_step(*c->sd);
}
void orc_set_image(void *p, int which, const double *in) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec == 4) { auto &a = c->sf->img[which]; for (size_t i = 0; i < a.size(); i++) a[i] = (float)in[i]; }
  else { auto &a = c->sd->img[which]; for (size_t i = 0; i < a.size(); i++) a[i] = in[i]; }
}

// Images on the global gridphy (x-fastest): which 0 Ip, 1 Ipp, 2 Ips, 3 Isp, 4 Iss.
void orc_get_image(void *p, int which, double *out) {
  OrcCtx *c = (OrcCtx *)p;
  if (c->prec == 4) { auto &a = c->sf->img[which]; for (size_t i = 0; i < a.size(); i++) out[i] = a[i]; }
  else { auto &a = c->sd->img[which]; for (size_t i = 0; i < a.size(); i++) out[i] = a[i]; }
}

// IMG/rtm3d_iamge_kernel.hpp:184-200 then :118-135, in the order of
// RTM/rtm3d_blend_elastic_model_mpi_meat.hpp:123-124, on one subdomain box of
// the global gridphy (i0..i0+ni etc.).  Outputs for the box, x-fastest:
// out_img[4][box] compensated Ipp,Ips,Isp,Iss and out_lap[4][box]
// Laplacians of the uncompensated images (zero on the 2-cell border).
void orc_finalize_box(void *p, int i0, int j0, int k0, int ni, int nj, int nk, double *out_img, double *out_lap) {
  OrcCtx *c = (OrcCtx *)p;
  auto run = [&](auto &s) {
    using T = typename std::remove_reference<decltype(s.lam[0])>::type;
    const int nx = s.d.nx, ny = s.d.ny;
    size_t nb = (size_t)ni * nj * nk;
    auto gid = [&](int i, int j, int k) { return (size_t)(i + i0) + (size_t)nx * ((size_t)(j + j0) + (size_t)ny * (k + k0)); };
    for (int w = 0; w < 4; w++) {
      const std::vector<T> &I = s.img[w + 1];
      for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
          for (int i = 0; i < ni; i++) {
            size_t q = (size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * k);
            T lap = 0;
            if (i >= 2 && i <= ni - 3 && j >= 2 && j <= nj - 3 && k >= 2 && k <= nk - 3) {
              lap = I[gid(i + 1, j, k)] + I[gid(i - 1, j, k)] + I[gid(i, j + 1, k)] + I[gid(i, j - 1, k)] +
                    I[gid(i, j, k + 1)] + I[gid(i, j, k - 1)] - T(6) * I[gid(i, j, k)];
            }
            out_lap[(size_t)w * nb + q] = (double)lap;
            T ip = s.img[0][gid(i, j, k)];
            T v = I[gid(i, j, k)];
            if ((double)ip > 3.4e-38) v = v / ip;  // double literal compare as in the reference
            out_img[(size_t)w * nb + q] = (double)v;
          }
    }
  };
  if (c->prec == 4) run(*c->sf); else run(*c->sd);
}

}  // extern "C"
