// Host side of the MI355X elastic FDTD / RTM engine: decomposition of the
// global model into subdomains that all live on one GPU, model extension,
// CPML coefficients, CMM halo and VOWR descriptor tables, the time-loop
// drivers and the C ABI declared in include/ddx.h.
//
// Device layout.  Each subdomain stores every field on its local extended
// grid (gridext) plus 4 ghost layers on each side that faces a neighbour
// subdomain; x is fastest and the x pitch is padded so that each row of the
// physical region starts on a 128-byte boundary.  Ghost layers are filled by
// the CMM exchange after every half step: one kernel per batch copies every
// neighbour's boundary layers straight into the receiver's ghosts (the
// contiguous message block of an MPI exchange is not needed on one device).
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ddx.h"
#include "ddx_internal.h"

namespace ddx {

static const double kPi = 3.14159265358979323846264338327950288;
static thread_local std::string g_create_error;

// ---------------------------------------------------------------------------
// Geometry
// ---------------------------------------------------------------------------
// Axis order everywhere: 0 x, 1 y, 2 z.  Side order of the reference
// (top, bottom, front, back, right, left) maps to (z lo, z hi, y lo, y hi,
// x lo, x hi).
static const int kSideAxis[6] = {2, 2, 1, 1, 0, 0};
static const int kSideHi[6] = {0, 1, 0, 1, 0, 1};

// Extra floats between consecutive fields of a subdomain, so the fields'
// base addresses do not all fall on the same HBM channel (a multiple of
// 256 B that is not a multiple of 4 KB).
#ifndef DDX_FIELD_PAD
#define DDX_FIELD_PAD 0
#endif
static constexpr long kFieldPad = DDX_FIELD_PAD;
static_assert(kFieldPad % 32 == 0, "fields start on 32-element boundaries (aligned pair accesses)");

// FLD/jarvis_3_device_field_meat.hpp:9-22 (get_gap_num).
static int gap_num(float l, float r, float d) {
  float v = 1.0f / d;
  int nl = (l > 0) ? (int)(l * v + 0.125f) : (int)(l * v - 0.125f);
  int nr = (r > 0) ? (int)(r * v + 0.125f) : (int)(r * v - 0.125f);
  return nr - nl;
}

struct Geom {
  int rank;
  int ci[3];            // index of the subdomain in the decomposition
  int pn[3], poff[3];   // gridphy size and offset inside the global gridphy
  int en[3];            // local gridext size
  int mlo[3], mhi[3];   // margins (pml_num on global sides, 0 toward neighbours)
  int nlo[3], nhi[3];   // neighbour ranks (-1: global boundary)
  int glo[3], ghi[3];   // ghost widths
  int pad_x;
  long sy, sz, nstore, origin;
  long fpitch;          // distance between consecutive fields of a subdomain (>= nstore)
  float lphy[3], rphy[3], lext[3], rext[3];
  long ext_goff[3];     // offset of local gridext inside the global gridext
};

// BASE/fdtd3d_base_3_geomodel_meat.hpp:579-704 (set_gridphy, set_gridext) and
// OP/jarvis_2_mpicu_halo_comm_operator_meat.hpp:26-101 (mpiFrame rank
// layout: x fastest).
static bool make_geom(const ddx_desc &d, int rank, Geom &g, std::string &err) {
  const int nd[3] = {d.x_divide_num, d.y_divide_num, d.z_divide_num};
  const int gn[3] = {d.nx, d.ny, d.nz};
  const float dd[3] = {d.dx, d.dy, d.dz};
  const int m = d.pml_num;
  const int nsub = nd[0] * nd[1] * nd[2];
  if (rank < 0 || rank >= nsub) { err = "rank out of range"; return false; }
  g.rank = rank;
  g.ci[0] = rank % (nd[0] * nd[1]) % nd[0];
  g.ci[1] = rank % (nd[0] * nd[1]) / nd[0];
  g.ci[2] = rank / (nd[0] * nd[1]);
  const int stride[3] = {1, nd[0], nd[0] * nd[1]};
  for (int a = 0; a < 3; a++) {
    const int ext = gn[a] + 2 * m;
    const int i = g.ci[a], n = nd[a];
    int pn;
    if (n == 1) pn = gn[a];
    else if (i == 0) pn = ext / n - m;
    else if (i < n - 1) pn = ext / n;
    else pn = ext / n + ext % n - m;
    g.pn[a] = pn;
    g.poff[a] = i == 0 ? 0 : i * (ext / n) - m;
    g.lphy[a] = i == 0 ? 0.f : (float)(i * (ext / n)) * dd[a] - (float)m * dd[a];
    g.rphy[a] = g.lphy[a] + (float)(pn - 1) * dd[a];
    g.nlo[a] = i > 0 ? rank - stride[a] : -1;
    g.nhi[a] = i < n - 1 ? rank + stride[a] : -1;
    g.mlo[a] = g.nlo[a] < 0 ? m : 0;
    g.mhi[a] = g.nhi[a] < 0 ? m : 0;
    g.en[a] = pn + g.mlo[a] + g.mhi[a];
    g.lext[a] = -(float)g.mlo[a] * dd[a] + g.lphy[a];
    g.rext[a] = g.lext[a] + (float)(g.en[a] - 1) * dd[a];
    g.glo[a] = g.nlo[a] >= 0 ? 4 : 0;
    g.ghi[a] = g.nhi[a] >= 0 ? 4 : 0;
    g.ext_goff[a] = (long)g.poff[a] + m - g.mlo[a];
    if (pn < 8) { err = "subdomain too thin (needs >= 8 physical cells per axis)"; return false; }
  }
  // the imaging kernels index a subdomain's gridphy and the copy kernels a
  // face box in 32 bits
  if ((long)g.pn[0] * g.pn[1] * g.pn[2] > INT_MAX / 2) {
    err = "subdomain gridphy too large (more than 2^30 cells); use a finer decomposition";
    return false;
  }
  // storage: align the first physical x cell of every row to 32 floats
  g.pad_x = g.glo[0] + ((32 - (g.glo[0] + g.mlo[0]) % 32) % 32);
  // (+2: the paired bricks read up to 5 cells past an odd-width gridphy row)
  const long wx = g.pad_x + g.en[0] + g.ghi[0] + 2;
  g.sy = (wx + 31) / 32 * 32;
  g.sz = g.sy * (g.glo[1] + g.en[1] + g.ghi[1]);
  g.nstore = g.sz * (g.glo[2] + g.en[2] + g.ghi[2]);
  g.fpitch = g.nstore + kFieldPad;
  g.origin = g.pad_x + g.sy * g.glo[1] + g.sz * g.glo[2];
  return true;
}

// Widths of the CMM halo each subdomain receives, per field and side
// (top, bottom, front, back, right, left), F = phy_fdorder_half = 4
// (PHY/fdtd3d_phy_mpi_1_host_meat.hpp:31-42, :251).
static const int kHaloW[10][6] = {
    {3, 4, 3, 4, 4, 3},  // vx
    {3, 4, 4, 3, 3, 4},  // vy
    {4, 3, 3, 4, 3, 4},  // vz
    {0, 0, 0, 0, 3, 4},  // sxx
    {0, 0, 3, 4, 0, 0},  // syy
    {3, 4, 0, 0, 0, 0},  // szz
    {0, 0, 4, 3, 4, 3},  // sxy
    {4, 3, 0, 0, 4, 3},  // sxz
    {4, 3, 4, 3, 0, 0},  // syz
    {3, 4, 3, 4, 3, 4},  // sau
};
// VOWR ring widths: 1 -> R, 0 -> R-1, -1 -> none
// (BASE/fdtd3d_base_5_wave_recon_meat.hpp:19-30).
static const int kRingW[10][6] = {
    {0, 1, 0, 1, 1, 0},      // vx
    {0, 1, 1, 0, 0, 1},      // vy
    {1, 0, 0, 1, 0, 1},      // vz
    {-1, -1, -1, -1, 0, 1},  // sxx
    {-1, -1, 0, 1, -1, -1},  // syy
    {0, 1, -1, -1, -1, -1},  // szz
    {-1, -1, 1, 0, 1, 0},    // sxy
    {1, 0, -1, -1, 1, 0},    // sxz
    {1, 0, 1, 0, -1, -1},    // syz
    {0, 1, 0, 1, 0, 1},      // sau
};
static int ring_w(int f, int side, int R) {
  int t = kRingW[f][side];
  return t < 0 ? 0 : (t ? R : R - 1);
}
// Store order of the fields in a VOWR slice (elasticWaveRecon::initialize).
static const int kRingFields[10] = {VX, VY, VZ, SXX, SYY, SZZ, SAU, SXY, SXZ, SYZ};

// GEO/jarvis_device_geometry_bones.hpp:29-49 (cu_cal_diff_coeff).
static void fd_coef(int half, float *out) {
  for (int n = 0; n < half; n++) {
    double a1 = 1, a2 = 1, a3 = 1;
    for (int i = 0; i < half; i++) {
      if (i != n) a1 = a1 * (2 * i + 1) * (2 * i + 1);
      if (i < n) a2 = a2 * ((2 * n + 1) * (2 * n + 1) - (2 * i + 1) * (2 * i + 1));
      if (i > n) a3 = a3 * ((2 * i + 1) * (2 * i + 1) - (2 * n + 1) * (2 * n + 1));
    }
    a1 = a1 * std::pow(-1.0, n);
    out[n] = (float)(a1 / ((2 * n + 1) * a2 * a3));
  }
}

// CPML/fdtd3d_cpml_mpi_1_host_meat.hpp:147-620: damping profile of local
// extended index i on an axis with en cells, low margin mlo, high margin mhi
// (a margin of 0 means that side faces a neighbour and is not damped).
static bool cpml_axis(int i, int en, int mlo, int mhi, float h, float vp, float fm, float dt, float &a,
                      float &b, float &ah, float &bh) {
  a = 0.f; b = 1.f; ah = 0.f; bh = 1.f;
  bool lo = mlo > 0 && i >= 2 && i < mlo;
  bool hi = mhi > 0 && i >= en - mhi && i < en - 2;
  if (!lo && !hi) return false;
  const float R = 0.001f, DAMP = 7.0f;
  const float amax = (float)(kPi * (double)fm);
  const int mm = lo ? mlo : mhi;
  const float L = float(mm - 2);
  float temp = std::log10(1.0f / R) * DAMP * vp / (2.0f * L * h);
  float rf, rh;
  if (lo) { rf = (mm - i) / L; rh = (mm - i - 0.5f) / L; }
  else { rf = (i - (en - mm - 1) - 0.5f) / L; rh = (i - (en - mm - 1)) / L; }
  float df = (float)(temp * std::pow((double)rf, 2));
  float alf = amax * (1.0f - rf);
  float dh = (float)(temp * std::pow((double)rh, 2));
  float alh = amax * (1.0f - rh);
  if (alf < 0) alf = 0;
  if (alh < 0) alh = 0;
  float bf = 0.f, bhf = 0.f, af = 0.f, ahf = 0.f;
  if (df >= 0) {
    bf = std::exp(-(df + alf) * dt);
    bhf = std::exp(-(dh + alh) * dt);
  }
  if (std::fabs(df) >= 1e-8) {
    af = df * (bf - 1.0f) / (df + alf);
    ahf = dh * (bhf - 1.0f) / (dh + alh);
  }
  a = af; b = bf; ah = ahf; bh = bhf;
  return true;
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
template <class T>
struct DevTable {  // a descriptor table and its first-block index array on the device
  T *d = nullptr;
  int *starts = nullptr;
  int n = 0, nblk = 0;
  double bytes = 0;      // algorithmic bytes moved by one launch (copies, imaging)
  double cells = 0;      // cells covered by one launch (stencil tables)
  double psi_cells = 0;  // sum over cells of the damped axes in range (CPML tables)
  PmlOrder ord;          // XCD-partitioned block order (CPML tables; pieces owned by the context)
};

struct PhyTables {  // physical-region tables of one brick shape and model, per wavefield set
  Tiling t{0, 0, 0, 0};
  int model = 0;
  DevTable<PhyDesc> set[2];
};

struct Slab {
  int sub;
  int i0, j0, k0, ni, nj, nk, axes;
  int xlo, xhi, ylo, yhi, zlo, zhi;
  long ncell;
  void *psi[2] = {nullptr, nullptr};  // memory variables of each set (storage format)
  float *coef[2] = {nullptr, nullptr};  // CPML coefficients of the modelling / migration model
  long in_range[3] = {0, 0, 0};
};

// HIP-event timing of the kernel classes (0 velocity, 1 stress, 2 CPML
// velocity, 3 CPML stress, 4 CMM halo exchange, 5 imaging).  Event pairs
// come from a pool that persists across ddx_profile calls, so timing a
// launch creates no event once the pool is warm.
constexpr int kClasses = 6;
struct Profile {
  int kernel = -1;   // class named by the last ddx_profile (-1: every class, -2: none)
  int mask = 0;      // classes being timed, bit k for class k
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
  std::vector<int> cls;  // class of pool entry i, for the entries in use
  double bytes[kClasses] = {};
};

}  // namespace ddx

using namespace ddx;

// Blocks per XCD group of the z-blocked bricks' block order (measured best
// of 0, 1, 2, 4, 8, 16 on the 256x256x200 RTM shot).
static const int kBrickGroup = 8;

struct ddx_ctx {
  ddx_desc d;
  std::string err;
  hipStream_t st = nullptr;
  // Optional second stream for the CPML halves, so they overlap the
  // physical-region kernel of the same half step (fork/join by events).
  hipStream_t st_pml = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool pml_overlap = false;
  // Optional stream for the reverse (receiver-side) propagation of the RTM
  // reverse loop, so it overlaps the backward reconstruction.
  hipStream_t st_rev = nullptr;
  hipEvent_t ev_rev_fork = nullptr, ev_rev_join = nullptr;
  bool set_overlap = false;
  bool ring_oob = false;  // a boundary-store copy asked for a slice outside the store (none launched)
  int nsub = 0;
  std::vector<Geom> g;
  int gext[3];
  float fd[5][4];
  // Brick shape of the physical-region kernels per mode (index = mode),
  // the fastest of a sweep on MI355X for the 2x2x1 256x256x200 RTM shot:
  // deep z columns where the register budget allows, shallower ones for the
  // backward/reverse velocity kernels, which carry the sau arms as well.
  // (index 4: FWI mode, z-blocked bricks only)
  Tiling tiling_vel[kModes] = {{64, 4, 1, 8}, {64, 4, 1, 8}, {64, 4, 1, 2}, {64, 4, 1, 4}, {64, 4, 1, 4}};
  Tiling tiling_stress[kModes] = {{64, 4, 1, 2}, {64, 4, 1, 4}, {64, 4, 1, 4}, {64, 4, 1, 4}, {64, 4, 1, 2}};
  // device memory
  std::vector<void *> allocs;
  std::vector<size_t> alloc_bytes;  // size of each allocation in allocs
  double dev_bytes = 0;
  std::vector<char *> wave[2];   // per subdomain: NF fields, fpitch elements apart
  // storage format of wavefields, memory variables and the VOWR store, its
  // element size, and the log2 scales of the field classes (0 velocity,
  // 1 stress) with whether the caller fixed them
  int fmt = STORE_F32;
  int esz = 4;
  int lg[2] = {0, 0};
  bool lg_user[2] = {false, false};
  // VOWR store format and element size; with FP32 wavefields and a 16-bit
  // store, its class scales (log2) and whether the caller fixed them
  int ring_fmt = STORE_F32;
  int ring_esz = 4;
  int rlg[2] = {0, 0};
  bool rlg_user[2] = {false, false};
  float wmax = 0.f;                  // wavelet peak |w|
  float vp_min = 0.f, rhovp = 0.f;   // model: minimum vp, mean rho * vp
  // per model (0: modelling, 1: migration) and subdomain: dtr, lam, mu (fpitch apart)
  std::vector<float *> model[2];
  bool have_mig = false;         // RTM passes use model 1
  bool have_model = false, have_wavelet = false, have_geometry = false;
  float *wavelet = nullptr;
  std::vector<Slab> slabs;
  std::vector<PhyTables> phy_cache;  // physical-region tables per brick shape
  int pml_cz = 1;                    // planes per thread of the CPML kernels (1, 2, 4, 8)
  int pml_group = 32;                // XCD group size of the CPML block order (0: launch order)
  int pml_order = 0;                 // 1: the XCD-partitioned CPML block order (PmlOrder); 0: pml_group
  DevTable<SlabDesc> pml[2][2];  // [model][set]
  // CMM halo links: [set][batch] batch 0 vel, 1 stress, 2 stress + sau
  DevTable<HaloDesc> halo[2][3];
  // exchange mode (ddx_set_exchange): 0 a direct copy of every link into the
  // receiver's ghosts; 1 the CMM message blocks: every sender side packs the
  // boundary layers of the batch's fields into one contiguous block (fields
  // in push order), which the receiver unpacks into its ghosts
  // 2: pack on store -- the kernels that update boundary layers write the
  // receivers' ghosts themselves (MirrorDesc; needs DDX_MIRROR=1 at
  // creation, and is then the default); a half step whose physical-region
  // kernel has no such stores (the one-plane and LDS-tiled kernels) is
  // followed by the mode-0 copy
  int exch_mode = 0;
  bool have_mirror = false;              // mirror tables exist (DDX_MIRROR=1 at creation)
  MirrorDesc *d_mir[2] = {nullptr, nullptr};  // [set][subdomain]
  DevTable<CopyDesc> cmm_pack[2][3], cmm_unpack[2][3];  // [set][batch], built on first use of mode 1
  char *cmm_block[2][3] = {};                           // the blocks of a set and batch, link after link
  // snapshots of one field of set 0 during ddx_rtm_shot (gridphy of every
  // subdomain, rank after rank): [pass][index] device blocks
  int snap_field = -1, snap_every = 0, snap_count = 0;
  DevTable<CopyDesc> snap_tab[NF];
  long snap_cells = 0;
  void *snap_buf = nullptr;  // 2 * snap_count * snap_cells elements (storage format)
  // VOWR: [which] 0 all (save), 1 vel (inject), 2 stress (inject); set 0 only
  DevTable<CopyDesc> ring[3];
  long ring_slice = 0;
  char *store = nullptr;     // ring_slice elements per time slice (storage format)
  long store_slices = 0;
  // per-slice scaling of a 16-bit store over FP32 wavefields
  // (ddx_set_ring_scaling): the exponents of every slice and field, and the
  // running maxima of the slice being saved
  bool ring_dyn = false;
  int *store_exp = nullptr;   // [slices][10]
  unsigned *ring_max = nullptr;  // [10]
  // block-scaled store (STORE_BFP16): ring_runs exponents per slice
  long ring_runs = 0;
  signed char *store_bexp = nullptr;  // [slices][ring_runs]
  // imaging
  DevTable<ImgDesc> img;
  std::vector<void *> images;  // per subdomain 9 arrays
  // geometry
  int nshot = 0;
  std::vector<int> shot_sub;    // per shot: owning subdomain or -1
  std::vector<long> shot_off;
  int nrec = 0;
  std::vector<int> rec_rank, rec_input;
  int *d_rsub = nullptr;
  long *d_roff = nullptr;
  void **d_vptr[2] = {nullptr, nullptr};
  std::vector<float *> seis;    // per shot [3][nrec][ntime]: modelled / observed data
  std::vector<float *> syn;     // per shot [3][nrec][ntime]: FWI synthetic data
  double cell_updates = 0;
  Profile prof;
  // graph replay of the modelling time step (ddx_set_graphs): one captured
  // step, the time index read from d_it, which the step itself advances
  bool use_graph = false;
  int *d_it = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_shot = -1;
  long graph_version = 0, graph_built = -1;  // invalidated when tables, geometry or scales change
  // chunked shots: kind of the shot in progress (0 none, 1 modelling, 2 RTM),
  // its index and the next stage
  int chunk_kind = 0, chunk_shot = -1, chunk_next = 0;
  int *d_flag = nullptr;  // device flag of ddx_check_finite
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
#define HIPCHECK(ctx, call)                                                                     \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      (ctx)->err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call;            \
      return DDX_EHIP;                                                                          \
    }                                                                                           \
  } while (0)

static int dalloc(ddx_ctx *c, void **p, size_t bytes) {
  if (bytes == 0) bytes = 4;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    c->err = "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e);
    *p = nullptr;
    return DDX_ENOMEM;
  }
  e = hipMemsetAsync(*p, 0, bytes, c->st);
  if (e != hipSuccess) {
    c->err = std::string("hipMemset failed: ") + hipGetErrorString(e);
    return DDX_EHIP;
  }
  c->allocs.push_back(*p);
  c->alloc_bytes.push_back(bytes);
  c->dev_bytes += (double)bytes;
  return DDX_OK;
}

// Free one allocation of dalloc (the stream is synchronized first: kernels
// in flight may still read it) and set *p to null.
static void dfree(ddx_ctx *c, void **p) {
  if (!*p) return;
  (void)hipStreamSynchronize(c->st);
  for (size_t k = 0; k < c->allocs.size(); k++)
    if (c->allocs[k] == *p) {
      (void)hipFree(*p);
      c->dev_bytes -= (double)c->alloc_bytes[k];
      c->allocs.erase(c->allocs.begin() + (long)k);
      c->alloc_bytes.erase(c->alloc_bytes.begin() + (long)k);
      break;
    }
  *p = nullptr;
}

template <class T>
static int upload_table(ddx_ctx *c, DevTable<T> &t, std::vector<T> &h, const std::vector<int> &starts,
                        int nblk) {
  t.n = (int)h.size();
  t.nblk = nblk;
  if (t.n == 0) return DDX_OK;
  int rc = dalloc(c, (void **)&t.d, sizeof(T) * h.size());
  if (rc) return rc;
  rc = dalloc(c, (void **)&t.starts, sizeof(int) * starts.size());
  if (rc) return rc;
  HIPCHECK(c, hipMemcpyAsync(t.d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, c->st));
  HIPCHECK(c, hipMemcpyAsync(t.starts, starts.data(), sizeof(int) * starts.size(), hipMemcpyHostToDevice, c->st));
  HIPCHECK(c, hipStreamSynchronize(c->st));
  return DDX_OK;
}

// XCD-partitioned block order of a CPML table (PmlOrder): every slab is
// split into one part per XCD, by z chunks of cz planes when the slab is
// deeper than its y extent (the y slabs and the x slabs), else by plane
// tiles (the z slabs), so each part's z (resp. y) stencil halo crosses to
// another XCD only at the part's two ends.  Each XCD walks its parts chunk
// by chunk in table order.
static void pml_order_host(const std::vector<SlabDesc> &d, int cz, std::vector<PmlPiece> &pieces, PmlOrder &o) {
  std::vector<PmlPiece> per[8];
  o = PmlOrder{};
  for (int s = 0; s < (int)d.size(); s++) {
    const int T = pml_plane_tiles(d[s].ni, d[s].nj, cz);
    const int Z = (d[s].nk + pml_planes(cz) - 1) / pml_planes(cz);
    const bool by_chunk = Z >= 8 && d[s].nk > d[s].nj;
    for (int x = 0; x < 8; x++) {
      PmlPiece P{s, o.xn[x], 0, 0, T};
      long cnt;
      if (by_chunk) {
        P.p0 = x * Z / 8;
        cnt = (long)((x + 1) * Z / 8 - P.p0) * T;
      } else {
        P.t0 = (int)((long)x * T / 8);
        P.tcnt = (int)((long)(x + 1) * T / 8) - P.t0;
        cnt = (long)Z * P.tcnt;
      }
      if (cnt <= 0) continue;
      o.xn[x] += (int)cnt;
      per[x].push_back(P);
    }
  }
  pieces.clear();
  int mx = 0;
  for (int x = 0; x < 8; x++) {
    o.xoff[x] = (int)pieces.size();
    pieces.insert(pieces.end(), per[x].begin(), per[x].end());
    mx = std::max(mx, o.xn[x]);
  }
  o.xoff[8] = (int)pieces.size();
  o.grid = 8 * mx;
}

// The XCD order of a context's CPML table (pieces in device memory of the context).
static int upload_pml_order(ddx_ctx *c, DevTable<SlabDesc> &t, const std::vector<SlabDesc> &h) {
  std::vector<PmlPiece> pieces;
  pml_order_host(h, c->pml_cz, pieces, t.ord);
  t.ord.pieces = nullptr;
  if (pieces.empty()) return DDX_OK;
  PmlPiece *dp = nullptr;
  int rc = dalloc(c, (void **)&dp, sizeof(PmlPiece) * pieces.size());
  if (rc) return rc;
  HIPCHECK(c, hipMemcpyAsync(dp, pieces.data(), sizeof(PmlPiece) * pieces.size(), hipMemcpyHostToDevice, c->st));
  HIPCHECK(c, hipStreamSynchronize(c->st));
  t.ord.pieces = dp;
  return DDX_OK;
}

static inline void *fptr(ddx_ctx *c, int set, int s, int f) {
  return c->wave[set][s] + ((long)f * c->g[s].fpitch + c->g[s].origin) * c->esz;
}
static float scale_of(const ddx_ctx *c, int cls) { return std::ldexp(1.0f, c->lg[cls]); }
// class of a field: 0 velocity, 1 stress, -1 FP32-only (FWI accumulators)
static int field_class(int f) {
  if (f <= VZ || (f >= VXP && f <= VZP)) return 0;
  if (f >= SXX && f <= SAU) return 1;
  return -1;
}
// host conversions of the storage formats (round to nearest even)
static uint16_t to_f16(float x) {
  const _Float16 h = (_Float16)x;
  uint16_t u;
  memcpy(&u, &h, 2);
  return u;
}
static float from_f16(uint16_t u) {
  _Float16 h;
  memcpy(&h, &u, 2);
  return (float)h;
}
static uint16_t to_bf16(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float from_bf16(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float x;
  memcpy(&x, &u, 4);
  return x;
}
// stored element -> value (unscaled) and back, for host transfers
static float load_elem(const ddx_ctx *c, const char *p, long i, float inv_scale) {
  if (c->fmt == STORE_F32) { float v; memcpy(&v, p + 4 * i, 4); return v; }
  uint16_t u;
  memcpy(&u, p + 2 * i, 2);
  return (c->fmt == STORE_F16 ? from_f16(u) : from_bf16(u)) * inv_scale;
}
static void store_elem(const ddx_ctx *c, char *p, long i, float v, float scale) {
  if (c->fmt == STORE_F32) { memcpy(p + 4 * i, &v, 4); return; }
  const uint16_t u = c->fmt == STORE_F16 ? to_f16(v * scale) : to_bf16(v * scale);
  memcpy(p + 2 * i, &u, 2);
}

// profiling hooks around launches of one kernel class
struct ProfScope {
  ddx_ctx *c;
  bool on;
  hipStream_t s;
  hipEvent_t b = nullptr;
  ProfScope(ddx_ctx *c_, int kernel, double bytes, hipStream_t stream = nullptr)
      : c(c_), on((c_->prof.mask >> kernel) & 1), s(stream ? stream : c_->st) {
    if (!on) return;
    Profile &p = c->prof;
    const size_t i = p.cls.size();
    if (i == p.pool.size()) {
      hipEvent_t a0 = nullptr, b0 = nullptr;
      if (hipEventCreate(&a0) != hipSuccess || hipEventCreate(&b0) != hipSuccess) { on = false; return; }
      p.pool.push_back({a0, b0});
    }
    p.cls.push_back(kernel);
    b = p.pool[i].second;
    (void)hipEventRecord(p.pool[i].first, s);
    p.bytes[kernel] += bytes;
  }
  ~ProfScope() {
    if (on) (void)hipEventRecord(b, s);
  }
};

// ---------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------
static int build_tables(ddx_ctx *c);

extern "C" int ddx_create(const ddx_desc *desc, ddx_ctx **out) {
  if (!desc || !out) { g_create_error = "null argument"; return DDX_EINVAL; }
  *out = nullptr;
  const ddx_desc &d = *desc;
  if (d.nx < 8 || d.ny < 8 || d.nz < 8 || d.dx <= 0 || d.dy <= 0 || d.dz <= 0 || d.dt <= 0 || d.ntime < 1) {
    g_create_error = "invalid grid, spacing, dt or ntime";
    return DDX_EINVAL;
  }
  if (d.pml_num < 4) { g_create_error = "pml_num must be >= 4"; return DDX_EINVAL; }
  if (d.rec_width < 1 || d.rec_width > 4) { g_create_error = "rec_width must be in 1..4"; return DDX_EINVAL; }
  if (d.storage < DDX_STORE_FP32 || d.storage > DDX_STORE_BF16) {
    g_create_error = "storage must be DDX_STORE_FP32, DDX_STORE_FP16 or DDX_STORE_BF16";
    return DDX_EINVAL;
  }
  if (d.ring_storage < DDX_STORE_FP32 || d.ring_storage > DDX_STORE_BFP16 ||
      (d.storage != DDX_STORE_FP32 && d.ring_storage != DDX_STORE_FP32 && d.ring_storage != d.storage)) {
    g_create_error = "ring_storage must be 0, or DDX_STORE_FP16 / DDX_STORE_BF16 / DDX_STORE_BFP16 over FP32 storage";
    return DDX_EINVAL;
  }
  if (d.x_divide_num < 1 || d.y_divide_num < 1 || d.z_divide_num < 1 ||
      d.x_divide_num * d.y_divide_num * d.z_divide_num > 512) {
    g_create_error = "invalid decomposition";
    return DDX_EINVAL;
  }
  ddx_ctx *c = new ddx_ctx();
  c->d = d;
  c->fmt = d.storage;
  c->esz = d.storage == DDX_STORE_FP32 ? 4 : 2;
  c->ring_fmt = d.ring_storage == DDX_STORE_FP32 ? c->fmt : d.ring_storage;
  c->ring_esz = c->ring_fmt == STORE_F32 ? 4 : 2;
  c->nsub = d.x_divide_num * d.y_divide_num * d.z_divide_num;
  c->gext[0] = d.nx + 2 * d.pml_num;
  c->gext[1] = d.ny + 2 * d.pml_num;
  c->gext[2] = d.nz + 2 * d.pml_num;
  c->g.resize(c->nsub);
  for (int r = 0; r < c->nsub; r++) {
    std::string e;
    if (!make_geom(d, r, c->g[r], e)) {
      g_create_error = e;
      delete c;
      return DDX_EINVAL;
    }
  }
  hipError_t he = hipSetDevice(d.device);
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
  if (const char *e = getenv("DDX_PML_OVERLAP")) c->pml_overlap = atoi(e) != 0;
  if (const char *e = getenv("DDX_SET_OVERLAP")) c->set_overlap = atoi(e) != 0;
  if (he == hipSuccess && c->set_overlap) {
    he = hipStreamCreateWithFlags(&c->st_rev, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&c->ev_rev_fork, hipEventDisableTiming);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&c->ev_rev_join, hipEventDisableTiming);
  }
  if (he == hipSuccess && c->pml_overlap) {
    he = hipStreamCreateWithFlags(&c->st_pml, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  }
  if (he != hipSuccess) {
    g_create_error = std::string("HIP: ") + hipGetErrorString(he);
    delete c;
    return DDX_EHIP;
  }
  for (int h = 0; h <= 4; h++) {
    for (int r = 0; r < 4; r++) c->fd[h][r] = 0.f;
    if (h > 0) fd_coef(h, c->fd[h]);
  }
  set_fd_table(c->fd);
  // CPML work split, for sweeps (FP32 storage): 1, 2, 4, 8 planes per thread
  // with cached taps, or kPmlTiled + 2, 4, 8 (18, 20, 24) for LDS plane tiles
  if (const char *e = getenv("DDX_PML_CZ")) {
    const int v = atoi(e);
    const int p = pml_planes(v);
    if (c->fmt == STORE_F32 && (p == 1 || p == 2 || p == 4 || p == 8) && (v < kPmlTiled || p > 1)) c->pml_cz = v;
  }
  if (const char *e = getenv("DDX_PML_GROUP")) {  // XCD group size of the CPML blocks, for sweeps
    const int v = atoi(e);
    if (v >= 0 && v <= 65536) c->pml_group = v;
  }
  // 1: the XCD-partitioned order (measured: 17% less CPML traffic, 4% slower), for sweeps
  if (const char *e = getenv("DDX_PML_ORDER")) c->pml_order = atoi(e) != 0;
  {  // XCD group size of the brick blocks (process-wide); DDX_BRICK_GROUP for sweeps
    const char *e = getenv("DDX_BRICK_GROUP");
    const int v = e ? atoi(e) : kBrickGroup;
    set_brick_group(v >= 0 && v <= 65536 ? v : kBrickGroup);
  }
  if (const char *e = getenv("DDX_TILING")) {  // "TXxTYxCYxKZ[xFORM]", for tile sweeps
    Tiling t{0, 0, 0, 0};
    const int nf = sscanf(e, "%dx%dx%dx%dx%d", &t.tx, &t.ty, &t.cy, &t.kz, &t.px);
    if (t.px == 3) t.px = 1, t.xl = 1;
    if (nf >= 4 && ((t.kz == 0 && t.px == 1 && !t.xl && c->fmt == STORE_F32) ||
                    (t.kz > 0 && tiling_supported(t, c->fmt))))
      for (int m = 0; m < kModes; m++)
        if (m != FWI || (t.kz > 0 && t.cy > 0 && t.px == 1 && !t.xl)) c->tiling_vel[m] = c->tiling_stress[m] = t;
  }
  int rc = build_tables(c);
  if (rc) {
    g_create_error = c->err;
    ddx_destroy(c);
    return rc;
  }
  *out = c;
  return DDX_OK;
}

extern "C" void ddx_destroy(ddx_ctx *c) {
  if (!c) return;
  if (c->st) (void)hipStreamSynchronize(c->st);
  for (auto &p : c->prof.pool) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->st) (void)hipStreamDestroy(c->st);
  if (c->st_pml) (void)hipStreamDestroy(c->st_pml);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->st_rev) (void)hipStreamDestroy(c->st_rev);
  if (c->ev_rev_fork) (void)hipEventDestroy(c->ev_rev_fork);
  if (c->ev_rev_join) (void)hipEventDestroy(c->ev_rev_join);
  delete c;
}

extern "C" const char *ddx_last_error(const ddx_ctx *c) {
  return c ? c->err.c_str() : g_create_error.c_str();
}

// The pack-on-store tables (MirrorDesc) of every subdomain of a set: for each
// side with a neighbour, the neighbour's storage and the widths it takes
// per field (table H1, the receiver's opposite side).  Built only when
// DDX_MIRROR=1 at creation: the stores into the neighbours' ghosts sit in
// the bricks' plane loops, whose loads of the next plane may then not pass
// them (the compiler cannot rule out aliasing), and the bricks ran 1.6-2x
// slower than the copy exchange saves (measured on the 256x256x200 RTM
// shot: velocity 208 -> 341 us, stress 216 -> 421 us per launch).
static int build_mirrors(ddx_ctx *c) {
  const char *e = getenv("DDX_MIRROR");
  bool any_nb = false;
  for (int s = 0; s < c->nsub; s++)
    for (int a = 0; a < 3; a++) any_nb |= c->g[s].nlo[a] >= 0 || c->g[s].nhi[a] >= 0;
  c->have_mirror = any_nb && e && atoi(e) == 1;
  if (!c->have_mirror) return DDX_OK;
  for (int set = 0; set < 2; set++) {
    std::vector<MirrorDesc> h(c->nsub);
    for (int s = 0; s < c->nsub; s++) {
      const Geom &g = c->g[s];
      MirrorDesc &m = h[s];
      memset(&m, 0, sizeof m);
      for (int a = 0; a < 3; a++) m.en[a] = g.en[a];
      for (int side = 0; side < 6; side++) {
        const int a = kSideAxis[side], hi = kSideHi[side];
        const int r = hi ? g.nhi[a] : g.nlo[a];
        if (r < 0) continue;
        const Geom &gr = c->g[r];
        MirrorSide &ms = m.s[side];
        ms.base = (char *)fptr(c, set, r, 0);
        ms.fpitch = gr.fpitch;
        ms.sy = gr.sy;
        ms.sz = gr.sz;
        ms.shift = hi ? -g.en[a] : gr.en[a];
        for (int f = 0; f < 10; f++) ms.w[f] = kHaloW[f][side ^ 1];
      }
    }
    if (int rc = dalloc(c, (void **)&c->d_mir[set], sizeof(MirrorDesc) * h.size())) return rc;
    HIPCHECK(c, hipMemcpyAsync(c->d_mir[set], h.data(), sizeof(MirrorDesc) * h.size(), hipMemcpyHostToDevice, c->st));
    HIPCHECK(c, hipStreamSynchronize(c->st));
  }
  c->exch_mode = 2;
  return DDX_OK;
}
static const MirrorDesc *mirror_of(const ddx_ctx *c, int set, int s) {
  if (!c->have_mirror) return nullptr;
  for (int a = 0; a < 3; a++)
    if (c->g[s].nlo[a] >= 0 || c->g[s].nhi[a] >= 0) return c->d_mir[set] + s;
  return nullptr;
}

// Physical-region descriptor tables for one brick shape (one per set); only
// the first-block indices depend on the shape.  Built once per shape.
static int build_phy_tables(ddx_ctx *c, const Tiling &t, int model, PhyTables &out) {
  out.t = t;
  out.model = model;
  for (int set = 0; set < 2; set++) {
    std::vector<PhyDesc> h;
    std::vector<int> st;
    int blk = 0;
    double cells = 0;
    for (int s = 0; s < c->nsub; s++) {
      const Geom &g = c->g[s];
      PhyDesc p;
      p.dt = c->d.dt;
      for (int f = 0; f < NF; f++) p.f[f] = fptr(c, set, s, f);
      p.dtr = c->model[model][s] + g.origin;
      p.lam = c->model[model][s] + g.fpitch + g.origin;
      p.mu = c->model[model][s] + 2 * g.fpitch + g.origin;
      p.sy = g.sy; p.sz = g.sz;
      p.ox = g.mlo[0]; p.oy = g.mlo[1]; p.oz = g.mlo[2];
      p.nx = g.pn[0]; p.ny = g.pn[1]; p.nz = g.pn[2];
      p.gx = g.poff[0]; p.gy = g.poff[1]; p.gz = g.poff[2];
      p.gnx = c->d.nx; p.gny = c->d.ny; p.gnz = c->d.nz;
      p.R = c->d.rec_width;
      p.tiles_x = (g.pn[0] + kTileX - 1) / kTileX;
      p.tiles_y = (g.pn[1] + kTileY - 1) / kTileY;
      p.mir = mirror_of(c, set, s);
      p.blk0 = blk;
      st.push_back(blk);
      blk += t.kz > 0 ? tiling_blocks(t, g.pn[0], g.pn[1], g.pn[2]) : p.tiles_x * p.tiles_y * g.pn[2];
      cells += (double)g.pn[0] * g.pn[1] * g.pn[2];
      h.push_back(p);
    }
    int rc = upload_table(c, out.set[set], h, st, blk);
    if (rc) return rc;
    out.set[set].cells = cells;
  }
  return DDX_OK;
}

static bool same_tiling(const Tiling &a, const Tiling &b) {
  return a.tx == b.tx && a.ty == b.ty && a.cy == b.cy && a.kz == b.kz && a.px == b.px && a.xl == b.xl;
}

// Tables of shape t over model `model` when the cache holds them, else null.
static const PhyTables *cached_phy_tables(const ddx_ctx *c, const Tiling &t, int model) {
  for (const PhyTables &p : c->phy_cache)
    if (same_tiling(p.t, t) && p.model == model) return &p;
  return nullptr;
}

// Tables of shape t over model `model`, from the cache (built on first use).
static const PhyTables *phy_tables(ddx_ctx *c, const Tiling &t, int model) {
  for (const PhyTables &p : c->phy_cache)
    if (same_tiling(p.t, t) && p.model == model) return &p;
  PhyTables p;
  if (build_phy_tables(c, t, model, p)) return nullptr;
  c->phy_cache.push_back(p);
  return &c->phy_cache.back();
}

// Build the tables of every shape in use now, so no allocation happens in
// the time loop.
static int prepare_phy_tables(ddx_ctx *c, int model) {
  for (int m = 0; m < kModes; m++)
    if (!phy_tables(c, c->tiling_vel[m], model) || !phy_tables(c, c->tiling_stress[m], model)) return DDX_EHIP;
  return DDX_OK;
}

// Model used by a half step: the migration model for the RTM passes when one
// is set, else the modelling model.
static int model_for(const ddx_ctx *c, int mode) { return (rtm_mode(mode) && c->have_mig) ? 1 : 0; }

// Descriptor tables that do not depend on the model: wavefield storage,
// physical-region tiles, CMM halo links, VOWR rings, imaging.
static int build_tables(ddx_ctx *c) {
  int rc;
  const int S = c->nsub;
  for (int set = 0; set < 2; set++) {
    c->wave[set].resize(S);
    for (int s = 0; s < S; s++) {
      rc = dalloc(c, (void **)&c->wave[set][s], (size_t)c->esz * NF * c->g[s].fpitch);
      if (rc) return rc;
    }
  }
  c->model[0].resize(S);
  for (int s = 0; s < S; s++) {
    rc = dalloc(c, (void **)&c->model[0][s], sizeof(float) * 3 * c->g[s].fpitch);
    if (rc) return rc;
  }
  if ((rc = build_mirrors(c))) return rc;
  rc = prepare_phy_tables(c, 0);
  if (rc) return rc;
  // CMM halo links: for every receiver subdomain and side with a neighbour,
  // the neighbour's boundary layers of each field of the batch, in the
  // reference's push order (vx, vy, vz | sxx, syy, szz, sxy, sxz, syz, sau).
  static const int kBatch[3][8] = {{VX, VY, VZ, -1}, {SXX, SYY, SZZ, SXY, SXZ, SYZ, -1},
                                   {SXX, SYY, SZZ, SXY, SXZ, SYZ, SAU, -1}};
  for (int batch = 0; batch < 3; batch++) {
    for (int set = 0; set < 2; set++) {
      std::vector<HaloDesc> hh;
      std::vector<int> st;
      int blk = 0;
      long cnt_all = 0;
      for (int r = 0; r < S; r++) {
        const Geom &gr = c->g[r];
        for (int side = 0; side < 6; side++) {
          const int a = kSideAxis[side], hi = kSideHi[side];
          const int nb = hi ? gr.nhi[a] : gr.nlo[a];
          if (nb < 0) continue;
          const Geom &gs = c->g[nb];
          for (int q = 0; kBatch[batch][q] >= 0; q++) {
            const int f = kBatch[batch][q];
            const int w = kHaloW[f][side];
            if (w <= 0) continue;
            int lo_r[3] = {0, 0, 0}, lo_s[3] = {0, 0, 0}, n[3] = {gr.en[0], gr.en[1], gr.en[2]};
            n[a] = w;
            if (hi) { lo_r[a] = gr.en[a]; lo_s[a] = 0; }
            else { lo_r[a] = -w; lo_s[a] = gs.en[a] - w; }
            HaloDesc h;
            h.src = fptr(c, set, nb, f);
            h.ssy = gs.sy; h.ssz = gs.sz;
            h.si = lo_s[0]; h.sj = lo_s[1]; h.sk = lo_s[2];
            h.dst = fptr(c, set, r, f);
            h.dsy = gr.sy; h.dsz = gr.sz;
            h.di = lo_r[0]; h.dj = lo_r[1]; h.dk = lo_r[2];
            h.ni = n[0]; h.nj = n[1]; h.nk = n[2];
            h.blk0 = blk;
            st.push_back(blk);
            const long cnt = (long)n[0] * n[1] * n[2];
            blk += halo_blocks(n[0], n[1], n[2]);
            cnt_all += cnt;
            hh.push_back(h);
          }
        }
      }
      rc = upload_table(c, c->halo[set][batch], hh, st, blk);
      if (rc) return rc;
      c->halo[set][batch].bytes = 2.0 * c->esz * cnt_all;  // one read and one write per element
    }
  }
  // VOWR rings on the global sides of each subdomain's gridphy
  // (OP/jarvis_1_mpicu_halo_store_operator_meat.hpp:122-135, stored only on
  // sides without a neighbour, :137-217).
  {
    const int R = c->d.rec_width;
    std::vector<CopyDesc> h[3];
    std::vector<int> st[3];
    int blk[3] = {0, 0, 0};
    long off = 0;
    for (int s = 0; s < S; s++) {
      const Geom &g = c->g[s];
      for (int fi = 0; fi < 10; fi++) {
        const int f = kRingFields[fi];
        for (int side = 0; side < 6; side++) {
          const int a = kSideAxis[side], hi = kSideHi[side];
          if ((hi ? g.nhi[a] : g.nlo[a]) >= 0) continue;
          const int w = ring_w(f, side, R);
          if (w <= 0) continue;
          int lo[3] = {g.mlo[0], g.mlo[1], g.mlo[2]}, n[3] = {g.pn[0], g.pn[1], g.pn[2]};
          n[a] = w;
          lo[a] = hi ? g.mlo[a] + g.pn[a] : g.mlo[a] - w;
          CopyDesc p;
          p.field = fptr(c, 0, s, f);
          p.sy = g.sy; p.sz = g.sz;
          p.i0 = lo[0]; p.j0 = lo[1]; p.k0 = lo[2];
          p.ni = n[0]; p.nj = n[1]; p.nk = n[2];
          p.buf_off = off;
          p.bi = p.bj = p.bk = 0;
          p.bni = p.ni; p.bnj = p.nj;
          p.cls = f <= VZ ? 0 : 1;
          p.fid = fi;
          long cnt = (long)n[0] * n[1] * n[2];
          const int which = f <= VZ ? 1 : 2;
          for (int t : {0, which}) {
            p.blk0 = blk[t];
            st[t].push_back(blk[t]);
            blk[t] += (int)((cnt + kBlock - 1) / kBlock);
            h[t].push_back(p);
          }
          // pack on store: the injection also writes the part of the ring
          // box that lies in a send band of another axis into that
          // receiver's ghost layers
          for (int side2 = 0; c->have_mirror && side2 < 6; side2++) {
            const int a2 = kSideAxis[side2], hi2 = kSideHi[side2];
            const int r = hi2 ? g.nhi[a2] : g.nlo[a2];
            const int w2 = kHaloW[f][side2 ^ 1];
            if (r < 0 || a2 == a || w2 <= 0) continue;
            const int b0 = hi2 ? g.en[a2] - w2 : 0, b1 = hi2 ? g.en[a2] : w2;
            const int c0 = std::max(lo[a2], b0), c1 = std::min(lo[a2] + n[a2], b1);
            if (c1 <= c0) continue;
            const Geom &gr = c->g[r];
            CopyDesc m = p;
            int mlo[3] = {lo[0], lo[1], lo[2]}, mn[3] = {n[0], n[1], n[2]}, sub[3] = {0, 0, 0};
            mlo[a2] = c0;
            mn[a2] = c1 - c0;
            sub[a2] = c0 - lo[a2];
            mlo[a2] += hi2 ? -g.en[a2] : gr.en[a2];
            m.field = fptr(c, 0, r, f);
            m.sy = gr.sy; m.sz = gr.sz;
            m.i0 = mlo[0]; m.j0 = mlo[1]; m.k0 = mlo[2];
            m.ni = mn[0]; m.nj = mn[1]; m.nk = mn[2];
            m.bi = sub[0]; m.bj = sub[1]; m.bk = sub[2];
            m.bni = n[0]; m.bnj = n[1];
            m.blk0 = blk[which];
            st[which].push_back(blk[which]);
            blk[which] += (int)(((long)mn[0] * mn[1] * mn[2] + kBlock - 1) / kBlock);
            h[which].push_back(m);
          }
          // a block floating point store starts every box on a run
          off += c->ring_fmt == STORE_BFP16 ? (cnt + kRingBlock - 1) / kRingBlock * kRingBlock : cnt;
        }
      }
    }
    c->ring_slice = off;
    c->ring_runs = (off + kRingBlock - 1) / kRingBlock;
    // the block-scaled save takes a run as kRingBlock consecutive lanes: its
    // table must hold whole boxes starting on runs
    if (c->ring_fmt == STORE_BFP16)
      for (const CopyDesc &p : h[0])
        if (p.buf_off % kRingBlock || p.bi || p.bj || p.bk || p.bni != p.ni || p.bnj != p.nj) {
          c->err = "internal: a VOWR save box is not a whole box on a run";
          return DDX_EINVAL;
        }
    for (int t = 0; t < 3; t++) {
      rc = upload_table(c, c->ring[t], h[t], st[t], blk[t]);
      if (rc) return rc;
    }
  }
  // imaging
  {
    std::vector<ImgDesc> h;
    std::vector<int> st;
    int blk = 0;
    const size_t esz = c->d.image_fp64 ? sizeof(double) : sizeof(float);
    c->images.assign((size_t)S * 9, nullptr);
    for (int s = 0; s < S; s++) {
      const Geom &g = c->g[s];
      ImgDesc p;
      for (int q = 0; q < 6; q++) {
        const int f = q < 3 ? VX + q : VXP + (q - 3);
        p.b[q] = fptr(c, 0, s, f);
        p.r[q] = fptr(c, 1, s, f);
      }
      p.sy = g.sy; p.sz = g.sz;
      p.ox = g.mlo[0]; p.oy = g.mlo[1]; p.oz = g.mlo[2];
      p.nx = g.pn[0]; p.ny = g.pn[1]; p.nz = g.pn[2];
      const long np = (long)g.pn[0] * g.pn[1] * g.pn[2];
      for (int q = 0; q < 9; q++) {
        rc = dalloc(c, &c->images[(size_t)s * 9 + q], esz * np);
        if (rc) return rc;
        p.img[q] = c->images[(size_t)s * 9 + q];
      }
      p.blk0 = blk;
      st.push_back(blk);
      blk += (int)((np + kBlock - 1) / kBlock);
      h.push_back(p);
    }
    rc = upload_table(c, c->img, h, st, blk);
    if (rc) return rc;
    double np_all = 0;
    for (auto &g : c->g) np_all += (double)g.pn[0] * g.pn[1] * g.pn[2];
    c->img.bytes = np_all * (12.0 * 4 + 5.0 * 2 * esz);
  }
  // receiver pointer tables
  for (int set = 0; set < 2; set++) {
    std::vector<void *> h((size_t)S * 3);
    for (int s = 0; s < S; s++)
      for (int q = 0; q < 3; q++) h[(size_t)s * 3 + q] = fptr(c, set, s, VX + q);
    rc = dalloc(c, (void **)&c->d_vptr[set], sizeof(void *) * h.size());
    if (rc) return rc;
    HIPCHECK(c, hipMemcpyAsync(c->d_vptr[set], h.data(), sizeof(void *) * h.size(), hipMemcpyHostToDevice, c->st));
  }
  rc = dalloc(c, (void **)&c->wavelet, sizeof(float) * c->d.ntime);
  if (rc) return rc;
  rc = dalloc(c, (void **)&c->d_it, sizeof(int));
  if (rc) return rc;
  rc = dalloc(c, (void **)&c->d_flag, sizeof(int));
  if (rc) return rc;
  HIPCHECK(c, hipStreamSynchronize(c->st));
  return DDX_OK;
}

// BASE/fdtd3d_base_3_geomodel_meat.hpp:405-460 (model build) and :705-718
// (set_model_para), CPML/fdtd3d_cpml_mpi_1_host_meat.hpp:147-620 (damping).
// CPML slab boxes of one subdomain (BASE/fdtd3d_base_6_pml_meat.hpp:29-93;
// sub_pad :213-223): z slabs over ext x by ext y, y slabs over ext x by phy
// z, x slabs over phy y by phy z, the outer 2 layers excluded; calls
// f(i0, i1, j0, j1, k0, k1, damped axes).
template <class F>
static void for_each_slab(const Geom &g, F &&f) {
  const int xb0 = g.mlo[0] > 0 ? 2 : 0, xb1 = g.en[0] - (g.mhi[0] > 0 ? 2 : 0);
  const int yb0 = g.mlo[1] > 0 ? 2 : 0, yb1 = g.en[1] - (g.mhi[1] > 0 ? 2 : 0);
  const int hasx = (g.mlo[0] > 0 || g.mhi[0] > 0) ? 1 : 0;
  const int hasy = (g.mlo[1] > 0 || g.mhi[1] > 0) ? 2 : 0;
  const int pz0 = g.mlo[2], pz1 = g.mlo[2] + g.pn[2];
  const int py0 = g.mlo[1], py1 = g.mlo[1] + g.pn[1];
  if (g.mlo[2] > 0) f(xb0, xb1, yb0, yb1, 2, g.mlo[2], hasx | hasy | 4);
  if (g.mhi[2] > 0) f(xb0, xb1, yb0, yb1, g.en[2] - g.mhi[2], g.en[2] - 2, hasx | hasy | 4);
  if (g.mlo[1] > 0) f(xb0, xb1, 2, g.mlo[1], pz0, pz1, hasx | 2);
  if (g.mhi[1] > 0) f(xb0, xb1, g.en[1] - g.mhi[1], g.en[1] - 2, pz0, pz1, hasx | 2);
  if (g.mlo[0] > 0) f(2, g.mlo[0], py0, py1, pz0, pz1, 1);
  if (g.mhi[0] > 0) f(g.en[0] - g.mhi[0], g.en[0] - 2, py0, py1, pz0, pz1, 1);
}

// f(k) for k in [0, n) on up to 8 host threads (DDX_HOST_THREADS overrides):
// the per-plane host loops of model loading, whose iterations write disjoint
// cells.
template <class F>
static void parallel_for(int n, F &&f) {
  int nt = (int)std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
  if (const char *e = getenv("DDX_HOST_THREADS")) nt = std::max(1, std::min(64, atoi(e)));
  nt = std::min(nt, n);
  if (nt <= 1) {
    for (int k = 0; k < n; k++) f(k);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++)
    th.emplace_back([&] {
      for (int k; (k = next.fetch_add(1)) < n;) f(k);
    });
  for (auto &x : th) x.join();
}

// Upload model `which` (0 modelling, 1 migration): extension to gridext,
// dt/rho, lambda, mu and the CPML coefficients; builds the CPML slab list on
// the first call and the CPML tables of this model.
static int load_model(ddx_ctx *c, int which, const float *vp, const float *vs, const float *rho) {
  const ddx_desc &d = c->d;
  const int m = d.pml_num;
  // stability (PHY/fdtd3d_phy_mpi_1_host_meat.hpp:142-148; the reference
  // computes dt_cal but its check is commented out)
  float vmax = 0.f;
  const long np = (long)d.nx * d.ny * d.nz;
  for (long i = 0; i < np; i++) {
    vmax = std::max(vmax, vp[i]);
    vmax = std::max(vmax, vs[i]);
    if (!(vp[i] > 0.f) || !(rho[i] > 0.f) || !(vs[i] >= 0.f) || !std::isfinite(vp[i])) {
      c->err = "model has non-positive or non-finite values";
      return DDX_EINVAL;
    }
  }
  const float dtc = ddx_stable_dt(vmax, d.dx, d.dy, d.dz);
  if (d.dt > dtc) {
    char buf[256];
    snprintf(buf, sizeof buf, "dt=%g exceeds the stability limit %g of the 8th-order scheme (vmax=%g)", d.dt,
             dtc, vmax);
    if (d.cfl_policy == 0) { c->err = buf; return DDX_EUNSTABLE; }
    fprintf(stderr, "ddx warning: %s\n", buf);
  }
  auto phy_at = [&](const float *a, long gi, long gj, long gk) {
    long pi = std::min(std::max(gi - m, 0L), (long)d.nx - 1);
    long pj = std::min(std::max(gj - m, 0L), (long)d.ny - 1);
    long pk = std::min(std::max(gk - m, 0L), (long)d.nz - 1);
    return a[pi + (long)d.nx * (pj + (long)d.ny * pk)];
  };
  // free old slab storage is kept (slabs are rebuilt only once)
  const bool first = c->slabs.empty();
  bool too_big = false;
  for (int s = 0; s < c->nsub; s++) {
    const Geom &g = c->g[s];
    std::vector<float> host((size_t)3 * g.fpitch, 0.f);
    std::vector<float> evp((size_t)g.en[0] * g.en[1] * g.en[2]);
    parallel_for(g.en[2], [&](int k) {
      for (int j = 0; j < g.en[1]; j++)
        for (int i = 0; i < g.en[0]; i++) {
          const long gi = i + g.ext_goff[0], gj = j + g.ext_goff[1], gk = k + g.ext_goff[2];
          const float p = phy_at(vp, gi, gj, gk), q = phy_at(vs, gi, gj, gk), r = phy_at(rho, gi, gj, gk);
          const long o = g.origin + i + g.sy * j + g.sz * k;
          host[o] = d.dt / r;                                   // dt / rho
          host[g.fpitch + o] = r * (p * p - 2.0f * q * q);      // lambda
          host[2 * g.fpitch + o] = r * (q * q);                 // mu
          evp[(size_t)i + (size_t)g.en[0] * ((size_t)j + (size_t)g.en[1] * k)] = p;
        }
    });
    HIPCHECK(c, hipMemcpyAsync(c->model[which][s], host.data(), sizeof(float) * host.size(), hipMemcpyHostToDevice,
                               c->st));
    HIPCHECK(c, hipStreamSynchronize(c->st));
    // CPML slabs (BASE/fdtd3d_base_6_pml_meat.hpp:29-93; sub_pad :213-223)
    if (first) {
      auto add = [&](int i0, int i1, int j0, int j1, int k0, int k1, int axes) {
        if (i1 <= i0 || j1 <= j0 || k1 <= k0) return;
        Slab sl;
        sl.sub = s;
        sl.i0 = i0; sl.j0 = j0; sl.k0 = k0;
        sl.ni = i1 - i0; sl.nj = j1 - j0; sl.nk = k1 - k0;
        sl.axes = axes;
        sl.xlo = g.mlo[0] > 0 ? g.mlo[0] : INT_MIN;
        sl.xhi = g.mhi[0] > 0 ? g.en[0] - g.mhi[0] : INT_MAX;
        sl.ylo = g.mlo[1] > 0 ? g.mlo[1] : INT_MIN;
        sl.yhi = g.mhi[1] > 0 ? g.en[1] - g.mhi[1] : INT_MAX;
        sl.zlo = g.mlo[2] > 0 ? g.mlo[2] : INT_MIN;
        sl.zhi = g.mhi[2] > 0 ? g.en[2] - g.mhi[2] : INT_MAX;
        sl.ncell = (long)sl.ni * sl.nj * sl.nk;
        if (sl.ncell > INT_MAX - kBlock) too_big = true;  // the CPML kernels index slabs in 32 bits
        c->slabs.push_back(sl);
      };
      for_each_slab(g, add);
      if (too_big) {
        c->err = "a CPML slab holds more than 2^31 cells";
        return DDX_EINVAL;
      }
    }
    // coefficients of this subdomain's slabs
    for (auto &sl : c->slabs) {
      if (sl.sub != s) continue;
      const int naxes = __builtin_popcount(sl.axes);
      if (!sl.coef[which]) {
        int rc = dalloc(c, (void **)&sl.coef[which], sizeof(float) * 4 * naxes * sl.ncell);
        if (rc) return rc;
      }
      if (!sl.psi[0]) {
        for (int set = 0; set < 2; set++) {
          int rc = dalloc(c, (void **)&sl.psi[set], (size_t)c->esz * 6 * naxes * sl.ncell);
          if (rc) return rc;
        }
      }
      std::vector<float> hc((size_t)4 * naxes * sl.ncell, 0.f);
      std::vector<long> in_plane((size_t)3 * sl.nk, 0);  // cells in each axis' damping range, per plane
      parallel_for(sl.nk, [&](int k) {
        long cidx = (long)sl.ni * sl.nj * k;
        for (int j = 0; j < sl.nj; j++)
          for (int i = 0; i < sl.ni; i++, cidx++) {
            const int ei = sl.i0 + i, ej = sl.j0 + j, ek = sl.k0 + k;
            const float v = evp[(size_t)ei + (size_t)g.en[0] * ((size_t)ej + (size_t)g.en[1] * ek)];
            int slot = 0;
            const int en_[3] = {ei, ej, ek};
            const float h_[3] = {d.dx, d.dy, d.dz};
            for (int a = 0; a < 3; a++) {
              if (!(sl.axes & (1 << a))) continue;
              float A, B, AH, BH;
              if (cpml_axis(en_[a], g.en[a], g.mlo[a], g.mhi[a], h_[a], v, d.fm, d.dt, A, B, AH, BH))
                in_plane[(size_t)3 * k + a]++;
              hc[(size_t)(slot * 4 + 0) * sl.ncell + cidx] = A;
              hc[(size_t)(slot * 4 + 1) * sl.ncell + cidx] = B;
              hc[(size_t)(slot * 4 + 2) * sl.ncell + cidx] = AH;
              hc[(size_t)(slot * 4 + 3) * sl.ncell + cidx] = BH;
              slot++;
            }
          }
      });
      sl.in_range[0] = sl.in_range[1] = sl.in_range[2] = 0;
      for (int k = 0; k < sl.nk; k++)
        for (int a = 0; a < 3; a++) sl.in_range[a] += in_plane[(size_t)3 * k + a];
      HIPCHECK(c, hipMemcpyAsync(sl.coef[which], hc.data(), sizeof(float) * hc.size(), hipMemcpyHostToDevice,
                                 c->st));
      HIPCHECK(c, hipStreamSynchronize(c->st));
    }
  }
  if (!c->pml[which][0].d) {
    for (int set = 0; set < 2; set++) {
      std::vector<SlabDesc> h;
      std::vector<int> st;
      int blk = 0;
      double cells = 0, psi_cells = 0;
      for (auto &sl : c->slabs) {
        const Geom &g = c->g[sl.sub];
        SlabDesc p;
        for (int f = 0; f < NF; f++) p.f[f] = fptr(c, set, sl.sub, f);
        p.dtr = c->model[which][sl.sub] + g.origin;
        p.lam = c->model[which][sl.sub] + g.fpitch + g.origin;
        p.mu = c->model[which][sl.sub] + 2 * g.fpitch + g.origin;
        p.sy = g.sy; p.sz = g.sz;
        p.i0 = sl.i0; p.j0 = sl.j0; p.k0 = sl.k0; p.ni = sl.ni; p.nj = sl.nj; p.nk = sl.nk;
        p.axes = sl.axes;
        p.xlo = sl.xlo; p.xhi = sl.xhi; p.ylo = sl.ylo; p.yhi = sl.yhi; p.zlo = sl.zlo; p.zhi = sl.zhi;
        p.psi = sl.psi[set];
        p.coef = sl.coef[which];
        p.mir = mirror_of(c, set, sl.sub);
        p.blk0 = blk;
        st.push_back(blk);
        blk += pml_blocks(sl.ni, sl.nj, sl.nk, c->pml_cz);
        h.push_back(p);
        cells += (double)sl.ncell;
        psi_cells += (double)(sl.in_range[0] + sl.in_range[1] + sl.in_range[2]);
      }
      int rc = upload_table(c, c->pml[which][set], h, st, blk);
      if (rc) return rc;
      if ((rc = upload_pml_order(c, c->pml[which][set], h))) return rc;
      c->pml[which][set].cells = cells;
      c->pml[which][set].psi_cells = psi_cells;
    }
  }
  return DDX_OK;
}

// Default FP16 scales.  The explosive source adds w to one cell per step and
// the cell sheds it at about vp dt / h per step, so the source cell's stress
// peaks near max|w| h / (vp_min dt); velocities stay below that stress over
// rho vp (measured: about 1/30 of it near the source).  Both estimates are
// put at 2^10, 64x under the FP16 maximum (65504), leaving room for the
// receiver-side field of the reverse pass; the normal range then reaches
// 2^-24 of the peaks.  BF16 keeps the FP32 exponent range: scale 1.  An
// FP16 VOWR store over FP32 wavefields takes the same scales (its values
// are forward wavefield values).
static void auto_scales(ddx_ctx *c) {
  c->graph_version++;
  const bool f16 = c->fmt == STORE_F16, ring16 = c->fmt == STORE_F32 && c->ring_fmt == STORE_F16;
  if (!(f16 || ring16) || !(c->wmax > 0.f) || !(c->vp_min > 0.f) || !(c->rhovp > 0.f)) return;
  const double h = std::min(c->d.dx, std::min(c->d.dy, c->d.dz));
  const double speak = (double)c->wmax * h / ((double)c->vp_min * c->d.dt);
  const double vpeak = speak / c->rhovp;
  int *lg = f16 ? c->lg : c->rlg;
  const bool *user = f16 ? c->lg_user : c->rlg_user;
  if (!user[1]) lg[1] = 10 - (int)std::ceil(std::log2(speak));
  if (!user[0]) lg[0] = 10 - (int)std::ceil(std::log2(vpeak));
}

// FP32 wavefields with an FP16 or BF16 VOWR store: the scales set and read
// are the store's (a block floating point store scales each run itself)
static bool ring_scaled(const ddx_ctx *c) {
  return c->fmt == STORE_F32 && (c->ring_fmt == STORE_F16 || c->ring_fmt == STORE_BF16);
}
static int ensure_ring_exp(ddx_ctx *c);

extern "C" int ddx_set_storage_scale(ddx_ctx *c, int field_class, int log2_scale) {
  if (!c || field_class < 0 || field_class > 1 || log2_scale < -60 || log2_scale > 60) return DDX_EINVAL;
  if (c->fmt == STORE_F32 && !ring_scaled(c) && log2_scale != 0) {
    c->err = "FP32 storage (and a block floating point store) has no scale";
    return DDX_EINVAL;
  }
  HIPCHECK(c, hipStreamSynchronize(c->st));
  if (ring_scaled(c)) {
    c->rlg[field_class] = log2_scale;
    c->rlg_user[field_class] = true;
  } else {
    c->lg[field_class] = log2_scale;
    c->lg_user[field_class] = true;
  }
  c->graph_version++;
  return DDX_OK;
}

// Per-slice, per-field scales of a 16-bit VOWR store over FP32 wavefields
// (mode 1) instead of the class scales (mode 0, the default).
extern "C" int ddx_set_ring_scaling(ddx_ctx *c, int mode) {
  if (!c || mode < 0 || mode > 1) return DDX_EINVAL;
  if (mode == 1 && !ring_scaled(c)) {
    c->err = "per-slice ring scaling needs FP32 wavefields with an FP16 or BF16 ring_storage";
    return DDX_EINVAL;
  }
  HIPCHECK(c, hipStreamSynchronize(c->st));
  c->ring_dyn = mode == 1;
  return ensure_ring_exp(c);
}

extern "C" int ddx_get_storage_scale(const ddx_ctx *c, int field_class, int *log2_scale) {
  if (!c || !log2_scale || field_class < 0 || field_class > 1) return DDX_EINVAL;
  *log2_scale = ring_scaled(c) ? c->rlg[field_class] : c->lg[field_class];
  return DDX_OK;
}

extern "C" int ddx_set_model(ddx_ctx *c, const float *vp, const float *vs, const float *rho) {
  if (!c || !vp || !vs || !rho) return DDX_EINVAL;
  int rc = load_model(c, 0, vp, vs, rho);
  if (rc) return rc;
  const long np = (long)c->d.nx * c->d.ny * c->d.nz;
  float vmin = vp[0];
  double sum = 0.0;
  for (long i = 0; i < np; i++) {
    vmin = std::min(vmin, vp[i]);
    sum += (double)rho[i] * vp[i];
  }
  c->vp_min = vmin;
  c->rhovp = (float)(sum / (double)np);
  auto_scales(c);
  c->have_model = true;
  return DDX_OK;
}

extern "C" int ddx_set_migration_model(ddx_ctx *c, const float *vp, const float *vs, const float *rho) {
  if (!c) return DDX_EINVAL;
  if (!vp && !vs && !rho) {
    c->have_mig = false;
    return DDX_OK;
  }
  if (!vp || !vs || !rho) return DDX_EINVAL;
  if (!c->have_model) {
    c->err = "set the modelling model (ddx_set_model) before the migration model";
    return DDX_ESTATE;
  }
  HIPCHECK(c, hipStreamSynchronize(c->st));
  if (c->model[1].empty()) {
    c->model[1].resize(c->nsub);
    for (int s = 0; s < c->nsub; s++) {
      int rc = dalloc(c, (void **)&c->model[1][s], sizeof(float) * 3 * c->g[s].fpitch);
      if (rc) return rc;
    }
  }
  int rc = load_model(c, 1, vp, vs, rho);
  if (rc) return rc;
  rc = prepare_phy_tables(c, 1);
  if (rc) return rc;
  c->have_mig = true;
  return DDX_OK;
}

extern "C" int ddx_set_wavelet(ddx_ctx *c, const float *w) {
  if (!c || !w) return DDX_EINVAL;
  HIPCHECK(c, hipMemcpyAsync(c->wavelet, w, sizeof(float) * c->d.ntime, hipMemcpyHostToDevice, c->st));
  HIPCHECK(c, hipStreamSynchronize(c->st));
  float wm = 0.f;
  for (int i = 0; i < c->d.ntime; i++) wm = std::max(wm, std::fabs(w[i]));
  c->wmax = wm;
  auto_scales(c);
  c->have_wavelet = true;
  return DDX_OK;
}

// UTIL/.jarvis_2_geometry/jarvis_device_geometry_meat.hpp:59-128: a receiver
// belongs to the subdomain whose gridphy contains it (inclusive, in metres);
// its node is get_gap_num from the subdomain's gridext origin.  A shot is
// injected by the subdomain whose gridext contains it
// (PHY/fdtd3d_phy_mpi_3_stress_kernel.hpp:1323-1334).
extern "C" int ddx_set_geometry(ddx_ctx *c, const float *shots, int nshot, const float *recv, int nrecv) {
  if (c) c->graph_version++;
  if (!c || nshot < 0 || nrecv < 0 || (nshot && !shots) || (nrecv && !recv)) return DDX_EINVAL;
  const float dd[3] = {c->d.dx, c->d.dy, c->d.dz};
  // a new geometry replaces the receiver tables and every shot's data
  dfree(c, (void **)&c->d_rsub);
  dfree(c, (void **)&c->d_roff);
  for (float *&q : c->seis) dfree(c, (void **)&q);
  for (float *&q : c->syn) dfree(c, (void **)&q);
  c->nshot = nshot;
  c->shot_sub.assign(nshot, -1);
  c->shot_off.assign(nshot, 0);
  for (int q = 0; q < nshot; q++) {
    const float *p = shots + 3 * q;
    for (int s = 0; s < c->nsub; s++) {
      const Geom &g = c->g[s];
      bool in = true;
      for (int a = 0; a < 3; a++) in = in && p[a] >= g.lext[a] && p[a] <= g.rext[a];
      if (!in) continue;
      int ii[3];
      for (int a = 0; a < 3; a++) ii[a] = gap_num(g.lext[a], p[a], dd[a]);
      c->shot_sub[q] = s;
      c->shot_off[q] = ii[0] + g.sy * ii[1] + g.sz * ii[2];
      break;
    }
  }
  std::vector<int> rsub;
  std::vector<long> roff;
  c->rec_rank.clear();
  c->rec_input.clear();
  for (int s = 0; s < c->nsub; s++) {
    const Geom &g = c->g[s];
    for (int r = 0; r < nrecv; r++) {
      const float *p = recv + 3 * r;
      bool in = true;
      for (int a = 0; a < 3; a++) in = in && p[a] >= g.lphy[a] && p[a] <= g.rphy[a];
      if (!in) continue;
      int ii[3];
      for (int a = 0; a < 3; a++) ii[a] = gap_num(g.lext[a], p[a], dd[a]);
      rsub.push_back(s);
      roff.push_back(ii[0] + g.sy * ii[1] + g.sz * ii[2]);
      c->rec_rank.push_back(s);
      c->rec_input.push_back(r);
    }
  }
  c->nrec = (int)rsub.size();
  if (c->nrec > 0) {
    int rc = dalloc(c, (void **)&c->d_rsub, sizeof(int) * c->nrec);
    if (rc) return rc;
    rc = dalloc(c, (void **)&c->d_roff, sizeof(long) * c->nrec);
    if (rc) return rc;
    HIPCHECK(c, hipMemcpyAsync(c->d_rsub, rsub.data(), sizeof(int) * c->nrec, hipMemcpyHostToDevice, c->st));
    HIPCHECK(c, hipMemcpyAsync(c->d_roff, roff.data(), sizeof(long) * c->nrec, hipMemcpyHostToDevice, c->st));
  }
  c->seis.assign(nshot, nullptr);
  c->syn.assign(nshot, nullptr);
  HIPCHECK(c, hipStreamSynchronize(c->st));
  c->have_geometry = true;
  return DDX_OK;
}

extern "C" int ddx_receivers(const ddx_ctx *c, int *nkept, int *rank_of_row, int *input_of_row) {
  if (!c) return DDX_EINVAL;
  if (nkept) *nkept = c->nrec;
  for (int r = 0; r < c->nrec; r++) {
    if (rank_of_row) rank_of_row[r] = c->rec_rank[r];
    if (input_of_row) input_of_row[r] = c->rec_input[r];
  }
  return DDX_OK;
}

// ---------------------------------------------------------------------------
// half steps
// ---------------------------------------------------------------------------
// Fields of each exchange batch in the reference's push order (vx, vy, vz |
// sxx, syy, szz, sxy, sxz, syz, sau).
static const int kBatchFields[3][8] = {{VX, VY, VZ, -1}, {SXX, SYY, SZZ, SXY, SXZ, SYZ, -1},
                                       {SXX, SYY, SZZ, SXY, SXZ, SYZ, SAU, -1}};

// CMM message blocks (OP/jarvis_2_mpicu_halo_comm_operator_meat.hpp:299-326
// extract_into_halo / inject_from_halo; one contiguous block per sender side
// and batch holding each field's boundary layers at its offset in push
// order, as the multiple-halo operator lays them out, :408-461).  Both ends
// live in the same HBM, so the sender's block is the receiver's: packing
// writes it, unpacking reads it into the receiver's ghost layers.
static int build_cmm(ddx_ctx *c) {
  if (c->cmm_pack[0][0].d || c->cmm_block[0][0]) return DDX_OK;
  const int S = c->nsub;
  for (int batch = 0; batch < 3; batch++)
    for (int set = 0; set < 2; set++) {
      std::vector<CopyDesc> hp, hu;
      std::vector<int> st;
      int blk = 0;
      long off = 0;
      for (int s = 0; s < S; s++) {
        const Geom &gs = c->g[s];
        for (int side = 0; side < 6; side++) {  // the sender's side toward the receiver
          const int a = kSideAxis[side], hi = kSideHi[side];
          const int r = hi ? gs.nhi[a] : gs.nlo[a];
          if (r < 0) continue;
          const Geom &gr = c->g[r];
          for (int q = 0; kBatchFields[batch][q] >= 0; q++) {
            const int f = kBatchFields[batch][q];
            const int w = kHaloW[f][side ^ 1];  // width the receiver takes on its opposite side
            if (w <= 0) continue;
            int lo_s[3] = {0, 0, 0}, lo_r[3] = {0, 0, 0}, n[3] = {gs.en[0], gs.en[1], gs.en[2]};
            n[a] = w;
            if (hi) { lo_s[a] = gs.en[a] - w; lo_r[a] = -w; }
            else { lo_s[a] = 0; lo_r[a] = gr.en[a]; }
            CopyDesc p{};
            p.ni = n[0]; p.nj = n[1]; p.nk = n[2];
            p.bni = n[0]; p.bnj = n[1];
            p.buf_off = off;
            p.cls = f <= VZ ? 0 : 1;
            p.fid = 0;
            p.blk0 = blk;
            CopyDesc u = p;
            p.field = fptr(c, set, s, f);
            p.sy = gs.sy; p.sz = gs.sz;
            p.i0 = lo_s[0]; p.j0 = lo_s[1]; p.k0 = lo_s[2];
            u.field = fptr(c, set, r, f);
            u.sy = gr.sy; u.sz = gr.sz;
            u.i0 = lo_r[0]; u.j0 = lo_r[1]; u.k0 = lo_r[2];
            const long cnt = (long)n[0] * n[1] * n[2];
            st.push_back(blk);
            blk += (int)((cnt + kBlock - 1) / kBlock);
            off += cnt;
            hp.push_back(p);
            hu.push_back(u);
          }
        }
      }
      if (int rc = upload_table(c, c->cmm_pack[set][batch], hp, st, blk)) return rc;
      if (int rc = upload_table(c, c->cmm_unpack[set][batch], hu, st, blk)) return rc;
      c->cmm_pack[set][batch].bytes = c->cmm_unpack[set][batch].bytes = 2.0 * c->esz * off;
      if (off > 0)
        if (int rc = dalloc(c, (void **)&c->cmm_block[set][batch], (size_t)c->esz * off)) return rc;
    }
  return DDX_OK;
}

// The CMM exchange of one batch of one set on stream st (exchange mode 0:
// one direct copy per link; 1: pack into the message blocks, then unpack).
static void exchange_on(ddx_ctx *c, int set, int batch, hipStream_t st) {
  if (c->exch_mode == 1) {
    const DevTable<CopyDesc> &p = c->cmm_pack[set][batch], &u = c->cmm_unpack[set][batch];
    if (p.nblk == 0) return;
    ProfScope ps(c, 4, p.bytes + u.bytes, st);
    launch_copy(+1, c->esz, p.d, p.starts, p.n, p.nblk, c->cmm_block[set][batch], st);
    launch_copy(-1, c->esz, u.d, u.starts, u.n, u.nblk, c->cmm_block[set][batch], st);
    return;
  }
  const DevTable<HaloDesc> &h = c->halo[set][batch];
  if (h.nblk == 0) return;
  ProfScope ps(c, 4, h.bytes, st);
  launch_halo(c->esz, h.d, h.starts, h.n, h.nblk, st);
}

static int exchange(ddx_ctx *c, int set, int batch) {
  exchange_on(c, set, batch, c->st);
  return DDX_OK;
}

// Whether a half step whose physical-region kernel ran with brick shape t
// has written the receivers' ghosts itself (pack on store): exchange mode
// 2 and a z-blocked brick (the one-plane kernels, kz = 0, and the LDS-tiled
// kernels, cy = 0, have no such stores).  The CPML kernels and the VOWR
// injection always store into the ghosts when the mirror tables exist.
static bool packed_on_store(const ddx_ctx *c, const Tiling &t) {
  return c->exch_mode == 2 && c->have_mirror && t.kz > 0 && t.cy > 0;
}
// The exchange that follows a half step of brick shape t: nothing after a
// packed store, else the copy of mode 0 or 1.
static void exchange_after(ddx_ctx *c, int set, int batch, const Tiling &t, hipStream_t st) {
  if (!packed_on_store(c, t)) exchange_on(c, set, batch, st);
}

// Algorithmic bytes per gridphy cell.  Velocity: 6 stress reads + 3 RMW of
// the velocities (field words, storage format) and dt/rho (FP32);
// backward/reverse add the sau read and 3 RMW of vxp..vzp; FWI adds 3 RMW of
// u and 3 writes of vrt.
static double phy_vel_bytes(const ddx_ctx *c, int mode) {
  const double fw = (mode == RTM_BACKWARD || mode == RTM_REVERSE) ? 19.0 : mode == FWI ? 21.0 : 12.0;
  return fw * c->esz + 4.0;
}
// Stress: 3 velocity reads + 6 RMW (+ sau RMW in the RTM modes) and lambda, mu.
static double phy_stress_bytes(const ddx_ctx *c, int mode) {
  return (rtm_mode(mode) ? 17.0 : 15.0) * c->esz + 8.0;
}
// Scale ratios folded into the kernels' constants (1 for FP32 storage).
static float ratio_vs(const ddx_ctx *c) { return std::ldexp(1.0f, c->lg[0] - c->lg[1]); }
static float ratio_sv(const ddx_ctx *c) { return std::ldexp(1.0f, c->lg[1] - c->lg[0]); }

static int check_mode(ddx_ctx *c, int mode) {
  if (mode == FWI && c->fmt != STORE_F32) {
    c->err = "the FWI mode needs FP32 storage";
    return DDX_EINVAL;
  }
  return DDX_OK;
}

// One launch of the physical-region kernel (0 velocity, 1 stress) of a set
// and mode with brick shape t.
static int phy_launch(ddx_ctx *c, int kernel, int set, int mode, const Tiling &t) {
  const PhyTables *pt = phy_tables(c, t, model_for(c, mode));
  if (!pt) return DDX_EHIP;
  const DevTable<PhyDesc> &ph = pt->set[set];
  bool launched;
  if (kernel == 0) {
    const float sign = mode == RTM_BACKWARD ? -1.f : 1.f;
    ProfScope ps(c, 0, phy_vel_bytes(c, mode) * ph.cells);
    launched = launch_vel_phy(t, mode, c->fmt, ph.d, ph.starts, ph.n, ph.nblk, c->d.dx, c->d.dy, c->d.dz,
                              sign * ratio_vs(c), c->d.nx, c->d.ny, c->d.nz, c->d.rec_width, c->st);
  } else {
    const float dt = mode == RTM_BACKWARD ? -c->d.dt : c->d.dt;
    ProfScope ps(c, 1, phy_stress_bytes(c, mode) * ph.cells);
    launched = launch_stress_phy(t, mode, c->fmt, ph.d, ph.starts, ph.n, ph.nblk, c->d.dx, c->d.dy, c->d.dz,
                                 dt * ratio_sv(c), c->d.nx, c->d.ny, c->d.nz, c->d.rec_width, c->st);
  }
  if (!launched) {
    c->err = "no physical-region kernel for this brick shape, storage format and mode";
    return DDX_EINVAL;
  }
  return DDX_OK;
}

static int vel_update(ddx_ctx *c, int set, int mode) {
  if (!c || set < 0 || set > 1 || mode < 0 || mode >= kModes) return DDX_EINVAL;
  if (!c->have_model) { c->err = "no model"; return DDX_ESTATE; }
  if (int rc = check_mode(c, mode)) return rc;
  const bool side = c->pml_overlap && mode != RTM_BACKWARD;
  hipStream_t pst = side ? c->st_pml : c->st;
  if (side) {
    HIPCHECK(c, hipEventRecord(c->ev_fork, c->st));
    HIPCHECK(c, hipStreamWaitEvent(c->st_pml, c->ev_fork, 0));
  }
  if (int rc = phy_launch(c, 0, set, mode, c->tiling_vel[mode])) return rc;
  if (mode != RTM_BACKWARD) {
    // 6 field reads + 3 RMW and dt/rho per cell, 3 memory variables RMW per
    // damped axis in range; the CPML coefficients count as 1D profiles (not
    // counted)
    const DevTable<SlabDesc> &pm = c->pml[model_for(c, mode)][set];
    ProfScope ps(c, 2, (12.0 * c->esz + 4.0) * pm.cells + 6.0 * c->esz * pm.psi_cells, pst);
    launch_vel_pml(c->pml_cz, c->fmt, pm.d, pm.starts, pm.n, pm.nblk, c->d.dx, c->d.dy, c->d.dz, ratio_vs(c),
                   c->pml_group, c->pml_order ? &pm.ord : nullptr, pst);
  }
  if (side) {
    HIPCHECK(c, hipEventRecord(c->ev_join, c->st_pml));
    HIPCHECK(c, hipStreamWaitEvent(c->st, c->ev_join, 0));
  }
  return DDX_OK;
}

static int stress_update(ddx_ctx *c, int set, int mode) {
  if (!c || set < 0 || set > 1 || mode < 0 || mode >= kModes) return DDX_EINVAL;
  if (!c->have_model) { c->err = "no model"; return DDX_ESTATE; }
  if (int rc = check_mode(c, mode)) return rc;
  const float dt = mode == RTM_BACKWARD ? -c->d.dt : c->d.dt;
  const bool side = c->pml_overlap && mode != RTM_BACKWARD;
  hipStream_t pst = side ? c->st_pml : c->st;
  if (side) {
    HIPCHECK(c, hipEventRecord(c->ev_fork, c->st));
    HIPCHECK(c, hipStreamWaitEvent(c->st_pml, c->ev_fork, 0));
  }
  if (int rc = phy_launch(c, 1, set, mode, c->tiling_stress[mode])) return rc;
  if (mode != RTM_BACKWARD) {
    // 3 field reads + 6 RMW (+ sau RMW) and lambda, mu per cell, 3 memory
    // variables RMW per damped axis in range
    const double w = rtm_mode(mode) ? 17.0 : 15.0;
    const DevTable<SlabDesc> &pm = c->pml[model_for(c, mode)][set];
    ProfScope ps(c, 3, (w * c->esz + 8.0) * pm.cells + 6.0 * c->esz * pm.psi_cells, pst);
    launch_stress_pml(c->pml_cz, rtm_mode(mode) ? 1 : 0, c->fmt, pm.d, pm.starts, pm.n, pm.nblk, c->d.dx, c->d.dy,
                      c->d.dz, dt * ratio_sv(c), c->pml_group, c->pml_order ? &pm.ord : nullptr, pst);
  }
  if (side) {
    HIPCHECK(c, hipEventRecord(c->ev_join, c->st_pml));
    HIPCHECK(c, hipStreamWaitEvent(c->st, c->ev_join, 0));
  }
  return DDX_OK;
}

// half steps followed by the CMM exchange of their batch
static int vel_half(ddx_ctx *c, int set, int mode) {
  int rc = vel_update(c, set, mode);
  if (rc) return rc;
  exchange_after(c, set, 0, c->tiling_vel[mode], c->st);
  return DDX_OK;
}
static int stress_half(ddx_ctx *c, int set, int mode) {
  int rc = stress_update(c, set, mode);
  if (rc) return rc;
  const bool sau = mode == RTM_BACKWARD || mode == RTM_REVERSE;
  exchange_after(c, set, sau ? 2 : 1, c->tiling_stress[mode], c->st);
  return DDX_OK;
}
extern "C" int ddx_step_vel(ddx_ctx *c, int set, int mode) { return vel_half(c, set, mode); }

extern "C" int ddx_record_synthetic(ddx_ctx *c, int set, int ishot, int it) {
  if (!c || set < 0 || set > 1 || ishot < 0 || ishot >= c->nshot || it < 0 || it >= c->d.ntime) return DDX_EINVAL;
  if (!c->syn[ishot]) {
    int rc = dalloc(c, (void **)&c->syn[ishot], sizeof(float) * 3 * (size_t)std::max(c->nrec, 1) * c->d.ntime);
    if (rc) return rc;
  }
  launch_recv_sample(c->fmt, c->d_vptr[set], c->d_rsub, c->d_roff, c->nrec, it, c->d.ntime, c->syn[ishot],
                     1.0f / scale_of(c, 0), nullptr, c->st);
  HIPCHECK(c, hipGetLastError());
  return DDX_OK;
}

extern "C" int ddx_get_synthetic(ddx_ctx *c, int ishot, float *out) {
  if (!c || !out || ishot < 0 || ishot >= c->nshot) return DDX_EINVAL;
  if (!c->syn[ishot]) { c->err = "shot has no synthetic data"; return DDX_ESTATE; }
  HIPCHECK(c, hipMemcpyAsync(out, c->syn[ishot], sizeof(float) * 3 * (size_t)c->nrec * c->d.ntime,
                             hipMemcpyDeviceToHost, c->st));
  HIPCHECK(c, hipStreamSynchronize(c->st));
  return DDX_OK;
}

extern "C" int ddx_inject_residual(ddx_ctx *c, int set, int ishot, int it) {
  if (!c || set < 0 || set > 1 || ishot < 0 || ishot >= c->nshot || it < 0 || it >= c->d.ntime) return DDX_EINVAL;
  if (int rc = check_mode(c, FWI)) return rc;
  if (!c->seis[ishot] || !c->syn[ishot]) {
    c->err = "residual needs observed data (ddx_model_shot / ddx_set_observed) and synthetic data";
    return DDX_ESTATE;
  }
  launch_recv_residual(c->d_vptr[set], c->d_rsub, c->d_roff, c->nrec, it, c->d.ntime, c->seis[ishot],
                       c->syn[ishot], c->st);
  HIPCHECK(c, hipGetLastError());
  return DDX_OK;
}
extern "C" int ddx_step_stress(ddx_ctx *c, int set, int mode) { return stress_half(c, set, mode); }

// Step hooks without the exchange, and the exchange, VOWR and receiver
// pieces on their own, so a caller can rebuild the reference's interleaving
// of RTM/rtm3d_blend_elastic_model_mpi_meat.hpp:64-113 (ddx_rtm_shot runs
// exactly these pieces in that order).
extern "C" int ddx_update_vel(ddx_ctx *c, int set, int mode) { return vel_update(c, set, mode); }
extern "C" int ddx_update_stress(ddx_ctx *c, int set, int mode) { return stress_update(c, set, mode); }
extern "C" int ddx_halo_exchange(ddx_ctx *c, int set, int batch) {
  if (!c || set < 0 || set > 1 || batch < 0 || batch > 2) return DDX_EINVAL;
  return exchange(c, set, batch);
}

extern "C" int ddx_source(ddx_ctx *c, int set, int ishot, int it, int sign) {
  if (!c || ishot < 0 || ishot >= c->nshot || it < 0 || it >= c->d.ntime || set < 0 || set > 1)
    return DDX_EINVAL;
  const int s = c->shot_sub[ishot];
  if (s < 0) return DDX_OK;  // shot outside every gridext: nothing injected
  launch_source(c->fmt, fptr(c, set, s, SXX), fptr(c, set, s, SYY), fptr(c, set, s, SZZ), c->shot_off[ishot],
                c->wavelet, it, sign > 0 ? 1.f : -1.f, c->d.source_mode, scale_of(c, 1), nullptr, c->st);
  return DDX_OK;
}

extern "C" int ddx_reset_wave(ddx_ctx *c, int set) {
  if (!c || set < 0 || set > 1) return DDX_EINVAL;
  for (int s = 0; s < c->nsub; s++)
    HIPCHECK(c, hipMemsetAsync(c->wave[set][s], 0, (size_t)c->esz * NF * c->g[s].fpitch, c->st));
  for (auto &sl : c->slabs)
    if (sl.psi[set])
      HIPCHECK(c, hipMemsetAsync(sl.psi[set], 0, (size_t)c->esz * 6 * __builtin_popcount(sl.axes) * sl.ncell, c->st));
  return DDX_OK;
}

static double ext_cells(const ddx_ctx *c) { return (double)c->gext[0] * c->gext[1] * c->gext[2]; }
static double phy_cells(const ddx_ctx *c) { return (double)c->d.nx * c->d.ny * c->d.nz; }

static int check_ready(ddx_ctx *c, int ishot) {
  if (!c->have_model || !c->have_wavelet || !c->have_geometry) {
    c->err = "model, wavelet and geometry must be set first";
    return DDX_ESTATE;
  }
  if (ishot < 0 || ishot >= c->nshot) { c->err = "shot index out of range"; return DDX_EINVAL; }
  return DDX_OK;
}

static int ensure_seis(ddx_ctx *c, int ishot) {
  if (!c->seis[ishot]) {
    return dalloc(c, (void **)&c->seis[ishot], sizeof(float) * 3 * (size_t)std::max(c->nrec, 1) * c->d.ntime);
  }
  return DDX_OK;
}

// One modelling time step (BLEND/fdtd3d_blend_elastic_model_mpi_meat.hpp:
// 55-66): velocity, source, stress, receivers.  itp: read the time index
// from this device int (graph capture) instead of it.
static int model_step(ddx_ctx *c, int ishot, int it, const int *itp) {
  int rc;
  if ((rc = vel_half(c, 0, PURE_FORWARD))) return rc;
  const int s = c->shot_sub[ishot];
  if (s >= 0)
    launch_source(c->fmt, fptr(c, 0, s, SXX), fptr(c, 0, s, SYY), fptr(c, 0, s, SZZ), c->shot_off[ishot], c->wavelet,
                  it, 1.f, c->d.source_mode, scale_of(c, 1), itp, c->st);
  if ((rc = stress_half(c, 0, PURE_FORWARD))) return rc;
  launch_recv_sample(c->fmt, c->d_vptr[0], c->d_rsub, c->d_roff, c->nrec, it, c->d.ntime, c->seis[ishot],
                     1.0f / scale_of(c, 0), itp, c->st);
  return DDX_OK;
}

// Capture one modelling step, with the step counter's advance, as a graph
// (rebuilt when the shot, the brick shapes, the geometry or the scales
// changed).  Every table the step uses exists before the capture, so no
// allocation happens inside it; the capture is thread-local, so other
// contexts keep allocating in other threads.
static int build_model_graph(ddx_ctx *c, int ishot) {
  if (c->gexec && c->graph_shot == ishot && c->graph_built == c->graph_version) return DDX_OK;
  if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
  if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
  if (!phy_tables(c, c->tiling_vel[PURE_FORWARD], 0) || !phy_tables(c, c->tiling_stress[PURE_FORWARD], 0))
    return DDX_EHIP;
  HIPCHECK(c, hipStreamSynchronize(c->st));
  HIPCHECK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  int rc = model_step(c, ishot, 0, c->d_it);
  if (!rc) launch_advance(c->d_it, 1, c->st);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->st, &g);
  if (rc || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (!rc) c->err = std::string("graph capture failed: ") + hipGetErrorString(e);
    return rc ? rc : DDX_EHIP;
  }
  const hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    c->gexec = nullptr;
    c->err = std::string("graph instantiation failed: ") + hipGetErrorString(ei);
    return DDX_EHIP;
  }
  c->graph = g;
  c->graph_shot = ishot;
  c->graph_built = c->graph_version;
  return DDX_OK;
}

// BLEND/fdtd3d_blend_elastic_model_mpi_meat.hpp:44-88 (mpi_sub_forward):
// velocity, source, stress, receivers at every step.
extern "C" int ddx_model_shot(ddx_ctx *c, int ishot) {
  if (!c) return DDX_EINVAL;
  int rc = check_ready(c, ishot);
  if (rc) return rc;
  rc = ensure_seis(c, ishot);
  if (rc) return rc;
  rc = ddx_reset_wave(c, 0);
  if (rc) return rc;
  const int nt = c->d.ntime;
  if (c->use_graph && c->prof.mask == 0 && !c->pml_overlap) {
    if ((rc = build_model_graph(c, ishot))) return rc;
    launch_set_int(c->d_it, 0, c->st);
    for (int it = 0; it < nt; it++) HIPCHECK(c, hipGraphLaunch(c->gexec, c->st));
  } else {
    for (int it = 0; it < nt; it++)
      if ((rc = model_step(c, ishot, it, nullptr))) return rc;
  }
  HIPCHECK(c, hipGetLastError());
  c->cell_updates += ext_cells(c) * nt;
  return DDX_OK;
}

// ---------------------------------------------------------------------------
// Fused batches of modelling shots.  The descriptor tables of every member
// context are concatenated (block offsets shifted) into tables of one
// launch, so a time step runs one velocity and one stress launch of the
// physical-region bricks and of the CPML kernels over all the members'
// cells; each member's halo exchange, source and receivers follow on the
// same stream, at the member's own time index.  The order of operations per
// cell is the single-shot one, so every member's seismograms equal its
// ddx_model_shot's.  ddx_model_batch runs n shots of equal length start to
// end; a ddx_batch takes members that join and leave at step boundaries
// (continuous batching).
// ---------------------------------------------------------------------------
template <class T>
struct HostTable {  // host copy of a device descriptor table
  std::vector<T> d;
  std::vector<int> s;
  int nblk = 0;
};

template <class T>
struct FusedTable {
  T *d = nullptr;
  int *starts = nullptr;
  int n = 0, nblk = 0;
  bool owned = true;  // false: a view into the fixed batch scratch
  PmlOrder ord;       // XCD-partitioned block order (CPML tables; pieces owned here)
  void reset() {
    if (owned && d) (void)hipFree(d);
    if (owned && starts) (void)hipFree(starts);
    if (owned && ord.pieces) (void)hipFree((void *)ord.pieces);
    d = nullptr;
    starts = nullptr;
    n = nblk = 0;
    ord = PmlOrder{};
  }
  ~FusedTable() { reset(); }
};

template <class T>
static hipError_t download_table(const DevTable<T> &t, HostTable<T> &h) {
  h.d.resize(t.n);
  h.s.resize(t.n);
  h.nblk = t.nblk;
  if (t.n == 0) return hipSuccess;
  hipError_t e = hipMemcpy(h.d.data(), t.d, sizeof(T) * t.n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(h.s.data(), t.starts, sizeof(int) * t.n, hipMemcpyDeviceToHost);
  return e;
}

// Tables of several members concatenated, block offsets shifted (host).
template <class T>
static void concat_tables(const std::vector<const HostTable<T> *> &parts, HostTable<T> &out) {
  out.d.clear();
  out.s.clear();
  int blk = 0;
  for (const HostTable<T> *p : parts) {
    for (size_t i = 0; i < p->d.size(); i++) {
      T x = p->d[i];
      x.blk0 += blk;
      out.d.push_back(x);
      out.s.push_back(p->s[i] + blk);
    }
    blk += p->nblk;
  }
  out.nblk = blk;
}

// A fused table in device memory of its own.
template <class T>
static hipError_t upload_owned(const HostTable<T> &h, FusedTable<T> &out) {
  out.reset();
  out.owned = true;
  out.n = (int)h.d.size();
  out.nblk = h.nblk;
  if (out.n == 0) return hipSuccess;
  hipError_t e = hipMalloc((void **)&out.d, sizeof(T) * h.d.size());
  if (e == hipSuccess) e = hipMalloc((void **)&out.starts, sizeof(int) * h.s.size());
  if (e == hipSuccess) e = hipMemcpy(out.d, h.d.data(), sizeof(T) * h.d.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(out.starts, h.s.data(), sizeof(int) * h.s.size(), hipMemcpyHostToDevice);
  return e;
}

// A fused CPML table with its XCD-partitioned block order.
static hipError_t upload_owned_pml(const HostTable<SlabDesc> &h, int cz, FusedTable<SlabDesc> &out) {
  hipError_t e = upload_owned(h, out);
  if (e != hipSuccess || h.d.empty()) return e;
  std::vector<PmlPiece> pieces;
  PmlOrder o;
  pml_order_host(h.d, cz, pieces, o);
  o.pieces = nullptr;
  if (!pieces.empty()) {
    PmlPiece *dp = nullptr;
    e = hipMalloc((void **)&dp, sizeof(PmlPiece) * pieces.size());
    if (e != hipSuccess) return e;
    o.pieces = dp;
    e = hipMemcpy(dp, pieces.data(), sizeof(PmlPiece) * pieces.size(), hipMemcpyHostToDevice);
  }
  out.ord = o;
  return e;
}

// One batch member: a context, its shot, its next time index, and host
// copies of its tables for the batch's brick shapes.
struct BatchMember {
  ddx_ctx *c = nullptr;
  int ishot = 0, it = 0;
  int join = 0;  // batch step of the shot's first time step
  HostTable<PhyDesc> hv, hs;
  HostTable<SlabDesc> hp;
};

// What members of one fused launch must share.
struct BatchKey {
  int device = 0, fmt = 0, lg0 = 0, lg1 = 0, pml_cz = 1, pml_group = 0, pml_order = 1;
  float dt = 0.f, dx = 0.f, dy = 0.f, dz = 0.f;
  Tiling tv{0, 0, 0, 0}, ts{0, 0, 0, 0};
};
static BatchKey batch_key(const ddx_ctx *c) {
  BatchKey k;
  k.device = c->d.device; k.fmt = c->fmt; k.lg0 = c->lg[0]; k.lg1 = c->lg[1];
  k.pml_cz = c->pml_cz; k.pml_group = c->pml_group; k.pml_order = c->pml_order;
  k.dt = c->d.dt; k.dx = c->d.dx; k.dy = c->d.dy; k.dz = c->d.dz;
  k.tv = c->tiling_vel[PURE_FORWARD];
  k.ts = c->tiling_stress[PURE_FORWARD];
  return k;
}
static bool batch_compatible(const BatchKey &k, const ddx_ctx *c) {
  return c->d.device == k.device && c->fmt == k.fmt && c->lg[0] == k.lg0 && c->lg[1] == k.lg1 &&
         c->pml_cz == k.pml_cz && c->d.dt == k.dt && c->d.dx == k.dx && c->d.dy == k.dy && c->d.dz == k.dz &&
         !c->pml_overlap;
}
static const char *kBatchMismatch =
    "batch members need one device, equal dt, spacing, storage format and scales, and no CPML stream overlap";

// Ready a context to join a batch: seismogram buffer, zeroed wavefields,
// tables of the batch's brick shapes (host copies).
// noalloc: allocate nothing -- the seismogram buffer and the tables of the
// batch's brick shapes must exist already (DDX_ESTATE otherwise).
static int batch_prepare(ddx_ctx *c, int ishot, const BatchKey &k, BatchMember &m, bool noalloc = false) {
  if (int rc = check_ready(c, ishot)) return rc;
  if (noalloc) {
    if (!c->seis[ishot]) {
      c->err = "no seismogram buffer for this shot yet (run ddx_model_shot or ddx_set_observed on it once)";
      return DDX_ESTATE;
    }
    if (!cached_phy_tables(c, k.tv, 0) || !cached_phy_tables(c, k.ts, 0)) {
      c->err = "the tables of the batch's brick shapes are not built on this context yet (run one shot with them)";
      return DDX_ESTATE;
    }
  }
  if (int rc = ensure_seis(c, ishot)) return rc;
  if (int rc = ddx_reset_wave(c, 0)) return rc;
  // build both shapes first: building one may move the other in the cache
  if (!phy_tables(c, k.tv, 0) || !phy_tables(c, k.ts, 0)) return DDX_EHIP;
  const PhyTables *a = phy_tables(c, k.tv, 0);
  const PhyTables *b = phy_tables(c, k.ts, 0);
  hipError_t e = download_table(a->set[0], m.hv);
  if (e == hipSuccess) e = download_table(b->set[0], m.hs);
  if (e == hipSuccess) e = download_table(c->pml[0][0], m.hp);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  if (e != hipSuccess) {
    c->err = std::string("HIP error: ") + hipGetErrorString(e) + " preparing a batch member";
    return DDX_EHIP;
  }
  m.c = c;
  m.ishot = ishot;
  m.it = 0;
  return DDX_OK;
}

struct FusedSet {
  FusedTable<PhyDesc> v, s;
  FusedTable<SlabDesc> p;
  FusedTable<MemberIO> io;  // sources and receivers of the members
};
// The fused tables of a batch on the host.
struct FusedHost {
  HostTable<PhyDesc> v, s;
  HostTable<SlabDesc> p;
  HostTable<MemberIO> io;
};
static void fuse_host(const std::vector<BatchMember> &ms, FusedHost &f) {
  std::vector<const HostTable<PhyDesc> *> pv, ps;
  std::vector<const HostTable<SlabDesc> *> pp;
  // one MemberIO part per member (its receiver blocks), so upload_fused
  // lays the members' receiver blocks out one after another
  std::vector<HostTable<MemberIO>> parts(ms.size());
  std::vector<const HostTable<MemberIO> *> pio;
  for (size_t q = 0; q < ms.size(); q++) {
    const BatchMember &m = ms[q];
    pv.push_back(&m.hv);
    ps.push_back(&m.hs);
    pp.push_back(&m.hp);
    ddx_ctx *c = m.c;
    MemberIO x{};
    const int s = c->shot_sub[m.ishot];
    if (s >= 0) {
      x.sxx = fptr(c, 0, s, SXX);
      x.syy = fptr(c, 0, s, SYY);
      x.szz = fptr(c, 0, s, SZZ);
      x.src_off = c->shot_off[m.ishot];
    }
    x.wavelet = c->wavelet;
    x.src_mode = c->d.source_mode;
    x.ss = scale_of(c, 1);
    x.vptr = c->d_vptr[0];
    x.rsub = c->d_rsub;
    x.roff = c->d_roff;
    x.nrec = c->nrec;
    x.seis = c->seis[m.ishot];
    x.inv_sv = 1.0f / scale_of(c, 0);
    x.sv = scale_of(c, 0);
    x.join = m.join;
    x.ntime = c->d.ntime;
    x.blk0 = 0;
    parts[q].d.push_back(x);
    parts[q].s.push_back(0);
    parts[q].nblk = (x.nrec + kBlock - 1) / kBlock;
    pio.push_back(&parts[q]);
  }
  concat_tables(pv, f.v);
  concat_tables(ps, f.s);
  concat_tables(pp, f.p);
  concat_tables(pio, f.io);
}
static hipError_t fuse_members(const std::vector<BatchMember> &ms, FusedSet &f) {
  FusedHost h;
  fuse_host(ms, h);
  hipError_t e = upload_owned(h.v, f.v);
  if (e == hipSuccess) e = upload_owned(h.s, f.s);
  if (e == hipSuccess) e = upload_owned_pml(h.p, ms.empty() ? 1 : ms[0].c->pml_cz, f.p);
  if (e == hipSuccess) e = upload_owned(h.io, f.io);
  return e;
}

// the exchange of a batch member after a fused half step of brick shape t
static void halo_on(ddx_ctx *c, int set, int batch, const Tiling &t, hipStream_t st) {
  exchange_after(c, set, batch, t, st);
}

// One modelling time step of every member (each at its own time index, the
// batch step gs minus its join step), in the order of model_step.
static bool fused_step(const BatchKey &k, const FusedSet &f, const std::vector<BatchMember> &ms, int gs,
                       hipStream_t st, const int *gsp = nullptr) {
  const float rvs = std::ldexp(1.0f, k.lg0 - k.lg1), rsv = std::ldexp(1.0f, k.lg1 - k.lg0);
  if (!launch_vel_phy(k.tv, PURE_FORWARD, k.fmt, f.v.d, f.v.starts, f.v.n, f.v.nblk, k.dx, k.dy, k.dz, rvs, 0, 0, 0,
                      0, st))
    return false;
  launch_vel_pml(k.pml_cz, k.fmt, f.p.d, f.p.starts, f.p.n, f.p.nblk, k.dx, k.dy, k.dz, rvs, k.pml_group,
                 k.pml_order ? &f.p.ord : nullptr, st);
  for (const BatchMember &m : ms) halo_on(m.c, 0, 0, k.tv, st);
  launch_source_batch(k.fmt, f.io.d, f.io.n, gs, gsp, st);
  if (!launch_stress_phy(k.ts, PURE_FORWARD, k.fmt, f.s.d, f.s.starts, f.s.n, f.s.nblk, k.dx, k.dy, k.dz, k.dt * rsv,
                         0, 0, 0, 0, st))
    return false;
  launch_stress_pml(k.pml_cz, 0, k.fmt, f.p.d, f.p.starts, f.p.n, f.p.nblk, k.dx, k.dy, k.dz, k.dt * rsv,
                    k.pml_group, k.pml_order ? &f.p.ord : nullptr, st);
  for (const BatchMember &m : ms) halo_on(m.c, 0, 1, k.ts, st);
  launch_recv_batch(k.fmt, f.io.d, f.io.starts, f.io.n, f.io.nblk, gs, gsp, st);
  return true;
}
static const char *kNoKernel = "no physical-region kernel for the batch's brick shape and storage format";

// One fused time step captured in a HIP graph with the advance of the
// device step counter gsp, so a step costs one graph launch instead of a
// launch per kernel (ddx_set_graphs on the batch's first context).
struct FusedGraph {
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  void reset() {
    if (x) (void)hipGraphExecDestroy(x);
    if (g) (void)hipGraphDestroy(g);
    x = nullptr;
    g = nullptr;
  }
  ~FusedGraph() { reset(); }
};
static int capture_fused(const BatchKey &k, const FusedSet &f, const std::vector<BatchMember> &ms, int *gsp,
                         hipStream_t st, FusedGraph &out, std::string &err) {
  out.reset();
  hipError_t e = hipStreamSynchronize(st);
  if (e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) { err = std::string("graph capture failed: ") + hipGetErrorString(e); return DDX_EHIP; }
  const bool ok = fused_step(k, f, ms, 0, st, gsp);
  if (ok) launch_advance(gsp, 1, st);
  e = hipStreamEndCapture(st, &out.g);
  if (!ok) { out.reset(); err = kNoKernel; return DDX_EINVAL; }
  if (e == hipSuccess) e = hipGraphInstantiate(&out.x, out.g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    out.reset();
    err = std::string("graph capture failed: ") + hipGetErrorString(e);
    return DDX_EHIP;
  }
  return DDX_OK;
}

// nsteps fused steps from batch step gs0, eagerly or by replaying one
// captured step (the counter gsp set to gs0 first).
static int run_fused(const BatchKey &k, const FusedSet &f, const std::vector<BatchMember> &ms, int gs0, int nsteps,
                     hipStream_t st, int *gsp, bool graphs, std::string &err) {
  if (!graphs) {
    for (int it = 0; it < nsteps; it++)
      if (!fused_step(k, f, ms, gs0 + it, st)) {
        (void)hipStreamSynchronize(st);
        err = kNoKernel;
        return DDX_EINVAL;
      }
    return DDX_OK;
  }
  FusedGraph gr;
  if (int rc = capture_fused(k, f, ms, gsp, st, gr, err)) return rc;
  launch_set_int(gsp, gs0, st);
  for (int it = 0; it < nsteps; it++) {
    const hipError_t e = hipGraphLaunch(gr.x, st);
    if (e != hipSuccess) { err = std::string("graph launch failed: ") + hipGetErrorString(e); return DDX_EHIP; }
  }
  return DDX_OK;
}

extern "C" int ddx_model_batch(ddx_ctx *const *ctxs, const int *ishots, int n) {
  if (!ctxs || !ishots || n < 1) return DDX_EINVAL;
  for (int q = 0; q < n; q++)
    if (!ctxs[q]) return DDX_EINVAL;
  ddx_ctx *c0 = ctxs[0];
  const BatchKey k = batch_key(c0);
  HIPCHECK(c0, hipSetDevice(k.device));
  std::vector<BatchMember> ms(n);
  for (int q = 0; q < n; q++) {
    ddx_ctx *c = ctxs[q];
    for (int r = 0; r < q; r++)
      if (ctxs[r] == c) { c0->err = "a context appears twice in the batch"; return DDX_EINVAL; }
    if (!batch_compatible(k, c) || c->d.ntime != c0->d.ntime) {
      c0->err = std::string(kBatchMismatch) + ", and equal ntime";
      return DDX_EINVAL;
    }
    if (int rc = batch_prepare(c, ishots[q], k, ms[q])) {
      if (q) c0->err = "batch member " + std::to_string(q) + ": " + c->err;
      return rc;
    }
  }
  FusedSet f;
  HIPCHECK(c0, fuse_members(ms, f));
  const int nt = c0->d.ntime;
  if (int rc = run_fused(k, f, ms, 0, nt, c0->st, c0->d_it, c0->use_graph && c0->prof.mask == 0, c0->err)) return rc;
  for (BatchMember &m : ms) m.it = nt;
  HIPCHECK(c0, hipGetLastError());
  HIPCHECK(c0, hipStreamSynchronize(c0->st));
  for (int q = 0; q < n; q++) ctxs[q]->cell_updates += ext_cells(ctxs[q]) * nt;
  return DDX_OK;
}

// ---------------------------------------------------------------------------
// ddx_model_batch_noalloc: the same fused batch with its descriptor tables
// (the only device memory ddx_model_batch allocates per call) in `scratch`,
// a fixed device array of kScratchBytes, instead of memory allocated per
// call.  Each device holds its own copy of the array; its address comes
// from hipGetSymbolAddress on that device, once, and is kept per device.
// Every call on a device rewrites the whole table area it uses, so nothing
// needs clearing; calls that overlap on one device race for the array.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = 4u << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[256];  // device address of `scratch` per device ordinal

// The device address of `scratch` on `device` (the current device).
static int scratch_of(int device, unsigned char **out, std::string &err) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= 256) { err = "device ordinal out of range"; return DDX_EINVAL; }
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
      return DDX_EHIP;
    }
    g_scratch_ptr[device] = p;
  }
  *out = (unsigned char *)g_scratch_ptr[device];
  return DDX_OK;
}

// Append v to the staging bytes at the next 256-byte boundary; returns its offset.
template <class T>
static size_t stage(std::vector<unsigned char> &buf, const std::vector<T> &v) {
  const size_t off = (buf.size() + 255) / 256 * 256;
  buf.resize(off + sizeof(T) * v.size());
  if (!v.empty()) memcpy(buf.data() + off, v.data(), sizeof(T) * v.size());
  return off;
}
template <class T>
static void view(FusedTable<T> &t, const HostTable<T> &h, unsigned char *dev, size_t od, size_t os) {
  t.owned = false;
  t.d = (T *)(dev + od);
  t.starts = (int *)(dev + os);
  t.n = (int)h.d.size();
  t.nblk = h.nblk;
}

// The fused tables of ms copied into scratch (dev) on stream st; f gets views.
static int fuse_into_scratch(const std::vector<BatchMember> &ms, FusedSet &f, unsigned char *dev, hipStream_t st,
                             std::string &err) {
  FusedHost h;
  fuse_host(ms, h);
  std::vector<PmlPiece> pieces;
  PmlOrder ord;
  pml_order_host(h.p.d, ms[0].c->pml_cz, pieces, ord);
  std::vector<unsigned char> buf;
  const size_t vd = stage(buf, h.v.d), vs = stage(buf, h.v.s);
  const size_t sd = stage(buf, h.s.d), ss = stage(buf, h.s.s);
  const size_t pd = stage(buf, h.p.d), ps = stage(buf, h.p.s), pp = stage(buf, pieces);
  const size_t id = stage(buf, h.io.d), is = stage(buf, h.io.s);
  if (buf.size() > kScratchBytes) {
    err = "the batch's descriptor tables need " + std::to_string(buf.size()) + " bytes, more than the " +
          std::to_string(kScratchBytes) + " of the fixed table area";
    return DDX_EINVAL;
  }
  hipError_t e = hipMemcpyAsync(dev, buf.data(), buf.size(), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);  // buf is pageable and goes out of scope
  if (e != hipSuccess) { err = std::string("HIP error: ") + hipGetErrorString(e); return DDX_EHIP; }
  view(f.v, h.v, dev, vd, vs);
  view(f.s, h.s, dev, sd, ss);
  view(f.p, h.p, dev, pd, ps);
  view(f.io, h.io, dev, id, is);
  f.p.ord = ord;
  f.p.ord.pieces = pieces.empty() ? nullptr : (const PmlPiece *)(dev + pp);
  return DDX_OK;
}

extern "C" int ddx_model_batch_noalloc(ddx_ctx *const *ctxs, const int *ishots, int n) {
  if (!ctxs || !ishots || n < 1) return DDX_EINVAL;
  for (int q = 0; q < n; q++)
    if (!ctxs[q]) return DDX_EINVAL;
  ddx_ctx *c0 = ctxs[0];
  const BatchKey k = batch_key(c0);
  HIPCHECK(c0, hipSetDevice(k.device));
  unsigned char *dev = nullptr;
  if (int rc = scratch_of(k.device, &dev, c0->err)) return rc;
  std::vector<BatchMember> ms(n);
  for (int q = 0; q < n; q++) {
    ddx_ctx *c = ctxs[q];
    for (int r = 0; r < q; r++)
      if (ctxs[r] == c) { c0->err = "a context appears twice in the batch"; return DDX_EINVAL; }
    if (!batch_compatible(k, c) || c->d.ntime != c0->d.ntime) {
      c0->err = std::string(kBatchMismatch) + ", and equal ntime";
      return DDX_EINVAL;
    }
    if (int rc = batch_prepare(c, ishots[q], k, ms[q], true)) {
      if (q) c0->err = "batch member " + std::to_string(q) + ": " + c->err;
      return rc;
    }
  }
  FusedSet f;
  if (int rc = fuse_into_scratch(ms, f, dev, c0->st, c0->err)) return rc;
  const int nt = c0->d.ntime;
  if (int rc = run_fused(k, f, ms, 0, nt, c0->st, nullptr, false, c0->err)) return rc;
  HIPCHECK(c0, hipGetLastError());
  HIPCHECK(c0, hipStreamSynchronize(c0->st));
  for (int q = 0; q < n; q++) ctxs[q]->cell_updates += ext_cells(ctxs[q]) * nt;
  return DDX_OK;
}

// ---------------------------------------------------------------------------
// Fused batches of RTM shots: n RTM shots of equal length run their forward
// passes with boundary save, then their reverse loops, together.  Every
// physical-region, CPML and imaging kernel of a half step is one launch over
// all members' cells (tables concatenated as for modelling batches; each
// descriptor carries its own grid size and ring width), sources and
// receiver injections one launch each; the CMM exchanges and the VOWR
// save/inject run per member.  The order of operations is ddx_rtm_shot's,
// so each member's images equal its own ddx_rtm_shot's.
// ---------------------------------------------------------------------------
static int rtm_begin(ddx_ctx *c, int ishot);
static void ring(ddx_ctx *c, int dir, int which, int it, hipStream_t st = nullptr);
static int ring_status(ddx_ctx *c, std::string &err);

struct RtmTables {  // the fused tables of one RTM batch
  FusedTable<PhyDesc> fv, fs, bv, bs, rv, rs;  // forward (set 0), backward (set 0), reverse (set 1)
  FusedTable<SlabDesc> fp, rp;                 // CPML of the forward and reverse sets
  FusedTable<MemberIO> src, inj;               // sources (set 0) and receiver injection (set 1)
  FusedTable<ImgDesc> img;
};

// joins (nullable): each member's MemberIO::join, with rev its direction
// (continuous batches); by default every member's time index is the step.
static int rtm_batch_tables(const std::vector<ddx_ctx *> &cs, const std::vector<int> &shots, RtmTables &T,
                            std::string &err, const std::vector<int> *joins = nullptr, int rev = 0) {
  ddx_ctx *c0 = cs[0];
  const int F = RTM_FORWARD, B = RTM_BACKWARD, R = RTM_REVERSE;
  const Tiling sh[6] = {c0->tiling_vel[F], c0->tiling_stress[F], c0->tiling_vel[B],
                        c0->tiling_stress[B], c0->tiling_vel[R], c0->tiling_stress[R]};
  const int mode_of[6] = {F, F, B, B, R, R}, set_of[6] = {0, 0, 0, 0, 1, 1};
  std::vector<HostTable<PhyDesc>> hp[6];
  std::vector<HostTable<SlabDesc>> hf(cs.size()), hr(cs.size());
  std::vector<HostTable<MemberIO>> hs(cs.size()), hi(cs.size());
  std::vector<HostTable<ImgDesc>> him(cs.size());
  for (auto &v : hp) v.resize(cs.size());
  hipError_t e = hipSuccess;
  for (size_t q = 0; q < cs.size() && e == hipSuccess; q++) {
    ddx_ctx *c = cs[q];
    // build every shape first: building one may move the others in the cache
    for (int k = 0; k < 6; k++)
      if (!phy_tables(c, sh[k], model_for(c, mode_of[k]))) { err = c->err; return DDX_EHIP; }
    for (int k = 0; k < 6 && e == hipSuccess; k++)
      e = download_table(phy_tables(c, sh[k], model_for(c, mode_of[k]))->set[set_of[k]], hp[k][q]);
    if (e == hipSuccess) e = download_table(c->pml[model_for(c, F)][0], hf[q]);
    if (e == hipSuccess) e = download_table(c->pml[model_for(c, R)][1], hr[q]);
    if (e == hipSuccess) e = download_table(c->img, him[q]);
    const int ishot = shots[q];
    MemberIO x{};
    const int s = c->shot_sub[ishot];
    if (s >= 0) {
      x.sxx = fptr(c, 0, s, SXX);
      x.syy = fptr(c, 0, s, SYY);
      x.szz = fptr(c, 0, s, SZZ);
      x.src_off = c->shot_off[ishot];
    }
    x.wavelet = c->wavelet;
    x.src_mode = c->d.source_mode;
    x.ss = scale_of(c, 1);
    x.rsub = c->d_rsub;
    x.roff = c->d_roff;
    x.nrec = c->nrec;
    x.seis = c->seis[ishot];
    x.inv_sv = 1.0f / scale_of(c, 0);
    x.sv = scale_of(c, 0);
    x.join = joins ? (*joins)[q] : 0;
    x.rev = rev;
    x.ntime = c->d.ntime;
    x.vptr = c->d_vptr[0];
    hs[q].d.push_back(x);
    hs[q].s.push_back(0);
    hs[q].nblk = 0;
    x.vptr = c->d_vptr[1];  // the observed data go into the reverse set
    hi[q].d.push_back(x);
    hi[q].s.push_back(0);
    hi[q].nblk = (x.nrec + kBlock - 1) / kBlock;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c0->st);
  auto ptrs = [](auto &v) {
    std::vector<const typename std::remove_reference<decltype(v[0])>::type *> out;
    for (auto &x : v) out.push_back(&x);
    return out;
  };
  FusedTable<PhyDesc> *pt[6] = {&T.fv, &T.fs, &T.bv, &T.bs, &T.rv, &T.rs};
  for (int k = 0; k < 6 && e == hipSuccess; k++) {
    HostTable<PhyDesc> h;
    concat_tables(ptrs(hp[k]), h);
    e = upload_owned(h, *pt[k]);
  }
  HostTable<SlabDesc> a, b;
  HostTable<MemberIO> m1, m2;
  HostTable<ImgDesc> im;
  concat_tables(ptrs(hf), a);
  concat_tables(ptrs(hr), b);
  concat_tables(ptrs(hs), m1);
  concat_tables(ptrs(hi), m2);
  concat_tables(ptrs(him), im);
  if (e == hipSuccess) e = upload_owned_pml(a, c0->pml_cz, T.fp);
  if (e == hipSuccess) e = upload_owned_pml(b, c0->pml_cz, T.rp);
  if (e == hipSuccess) e = upload_owned(m1, T.src);
  if (e == hipSuccess) e = upload_owned(m2, T.inj);
  if (e == hipSuccess) e = upload_owned(im, T.img);
  if (e != hipSuccess) { err = std::string("HIP: ") + hipGetErrorString(e); return DDX_EHIP; }
  return DDX_OK;
}

// One forward step with boundary save (rtm_forward_step) of every member of
// T at batch step gs; its[q]: member q's time index (the sources take theirs
// from their MemberIO).  c0: the member whose brick shapes T was built with.
static bool rtm_fused_forward(const BatchKey &k, const ddx_ctx *c0, const RtmTables &T,
                              const std::vector<ddx_ctx *> &cs, const std::vector<int> &its, int gs, hipStream_t st) {
  const int F = RTM_FORWARD;
  const float rvs = std::ldexp(1.0f, k.lg0 - k.lg1), rsv = std::ldexp(1.0f, k.lg1 - k.lg0);
  if (!launch_vel_phy(c0->tiling_vel[F], F, k.fmt, T.fv.d, T.fv.starts, T.fv.n, T.fv.nblk, k.dx, k.dy, k.dz, rvs, 0,
                      0, 0, 0, st))
    return false;
  launch_vel_pml(k.pml_cz, k.fmt, T.fp.d, T.fp.starts, T.fp.n, T.fp.nblk, k.dx, k.dy, k.dz, rvs, k.pml_group,
                 k.pml_order ? &T.fp.ord : nullptr, st);
  for (ddx_ctx *c : cs) halo_on(c, 0, 0, c0->tiling_vel[F], st);
  launch_source_batch(k.fmt, T.src.d, T.src.n, gs, nullptr, st, 1.f);
  if (!launch_stress_phy(c0->tiling_stress[F], F, k.fmt, T.fs.d, T.fs.starts, T.fs.n, T.fs.nblk, k.dx, k.dy, k.dz,
                         k.dt * rsv, 0, 0, 0, 0, st))
    return false;
  launch_stress_pml(k.pml_cz, 1, k.fmt, T.fp.d, T.fp.starts, T.fp.n, T.fp.nblk, k.dx, k.dy, k.dz, k.dt * rsv,
                    k.pml_group, k.pml_order ? &T.fp.ord : nullptr, st);
  for (ddx_ctx *c : cs) halo_on(c, 0, 1, c0->tiling_stress[F], st);
  // each member saves slices 0 .. its own ntime - 2 (members of a
  // continuous batch may differ in ntime)
  for (size_t q = 0; q < cs.size(); q++)
    if (its[q] <= cs[q]->d.ntime - 2) ring(cs[q], +1, 0, its[q], st);
  return true;
}

// One reverse iteration (rtm_reverse_step) of every member of T at batch
// step gs; its[q]: member q's time index.
static bool rtm_fused_reverse(const BatchKey &k, const ddx_ctx *c0, const RtmTables &T,
                              const std::vector<ddx_ctx *> &cs, const std::vector<int> &its, int gs, hipStream_t st) {
  const int B = RTM_BACKWARD, R = RTM_REVERSE;
  const float rvs = std::ldexp(1.0f, k.lg0 - k.lg1), rsv = std::ldexp(1.0f, k.lg1 - k.lg0);
  const float dx = k.dx, dy = k.dy, dz = k.dz, dt = k.dt;
  auto vel = [&](const FusedTable<PhyDesc> &t, int mode, float sign) {
    return launch_vel_phy(c0->tiling_vel[mode], mode, k.fmt, t.d, t.starts, t.n, t.nblk, dx, dy, dz, sign * rvs, 0,
                          0, 0, 0, st);
  };
  auto stress = [&](const FusedTable<PhyDesc> &t, int mode, float d) {
    return launch_stress_phy(c0->tiling_stress[mode], mode, k.fmt, t.d, t.starts, t.n, t.nblk, dx, dy, dz, d * rsv,
                             0, 0, 0, 0, st);
  };
  launch_source_batch(k.fmt, T.src.d, T.src.n, gs, nullptr, st, -1.f);
  launch_inject_batch(k.fmt, T.inj.d, T.inj.starts, T.inj.n, T.inj.nblk, gs, st);
  if (!stress(T.bs, B, -dt)) return false;
  for (size_t q = 0; q < cs.size(); q++)
    if (its[q] >= 1) ring(cs[q], -1, 2, its[q] - 1, st);
  for (ddx_ctx *c : cs) halo_on(c, 0, 2, c0->tiling_stress[B], st);
  if (!vel(T.rv, R, 1.f)) return false;
  launch_vel_pml(k.pml_cz, k.fmt, T.rp.d, T.rp.starts, T.rp.n, T.rp.nblk, dx, dy, dz, rvs, k.pml_group,
                 k.pml_order ? &T.rp.ord : nullptr, st);
  for (ddx_ctx *c : cs) halo_on(c, 1, 0, c0->tiling_vel[R], st);
  if (!vel(T.bv, B, -1.f)) return false;
  for (size_t q = 0; q < cs.size(); q++)
    if (its[q] >= 1) ring(cs[q], -1, 1, its[q] - 1, st);
  for (ddx_ctx *c : cs) halo_on(c, 0, 0, c0->tiling_vel[B], st);
  if (!stress(T.rs, R, dt)) return false;
  launch_stress_pml(k.pml_cz, 1, k.fmt, T.rp.d, T.rp.starts, T.rp.n, T.rp.nblk, dx, dy, dz, dt * rsv, k.pml_group,
                    k.pml_order ? &T.rp.ord : nullptr, st);
  for (ddx_ctx *c : cs) halo_on(c, 1, 2, c0->tiling_stress[R], st);
  launch_image(c0->d.image_fp64, k.fmt, T.img.d, T.img.starts, T.img.n, T.img.nblk, 1.0f / std::ldexp(1.0f, k.lg0),
               st);
  return true;
}

extern "C" int ddx_rtm_batch(ddx_ctx *const *ctxs, const int *ishots, int n) {
  if (!ctxs || !ishots || n < 1) return DDX_EINVAL;
  for (int q = 0; q < n; q++)
    if (!ctxs[q]) return DDX_EINVAL;
  ddx_ctx *c0 = ctxs[0];
  const BatchKey k = batch_key(c0);
  HIPCHECK(c0, hipSetDevice(k.device));
  std::vector<ddx_ctx *> cs(ctxs, ctxs + n);
  std::vector<int> shots(ishots, ishots + n);
  for (int q = 0; q < n; q++) {
    ddx_ctx *c = cs[q];
    for (int r = 0; r < q; r++)
      if (cs[r] == c) { c0->err = "a context appears twice in the batch"; return DDX_EINVAL; }
    if (!batch_compatible(k, c) || c->d.ntime != c0->d.ntime || c->d.image_fp64 != c0->d.image_fp64 ||
        c->set_overlap) {
      c0->err = std::string(kBatchMismatch) + ", equal ntime and image precision, and no reverse-set overlap";
      return DDX_EINVAL;
    }
    int rc = rtm_begin(c, shots[q]);
    if (!rc) rc = ddx_reset_wave(c, 1);
    if (!rc) rc = hipStreamSynchronize(c->st) == hipSuccess ? DDX_OK : DDX_EHIP;
    if (rc) {
      if (q) c0->err = "batch member " + std::to_string(q) + ": " + c->err;
      return rc;
    }
  }
  RtmTables T;
  if (int rc = rtm_batch_tables(cs, shots, T, c0->err)) return rc;
  const hipStream_t st = c0->st;
  const int nt = c0->d.ntime;
  bool ok = true;
  std::vector<int> its(cs.size());
  for (int it = 0; it < nt && ok; it++) {
    std::fill(its.begin(), its.end(), it);
    ok = rtm_fused_forward(k, c0, T, cs, its, it, st);
  }
  for (int it = nt - 1; it >= 0 && ok; it--) {
    std::fill(its.begin(), its.end(), it);
    ok = rtm_fused_reverse(k, c0, T, cs, its, it, st);
  }
  if (!ok) {
    (void)hipStreamSynchronize(st);
    c0->err = kNoKernel;
    return DDX_EINVAL;
  }
  for (ddx_ctx *c : cs)
    if (int rc = ring_status(c, c0->err)) return rc;
  HIPCHECK(c0, hipGetLastError());
  HIPCHECK(c0, hipStreamSynchronize(st));
  for (ddx_ctx *c : cs) c->cell_updates += (2.0 * ext_cells(c) + phy_cells(c)) * nt;
  return DDX_OK;
}

// Continuous batch: members join (ddx_batch_add) and leave when their shot
// has run all its steps (ddx_batch_run returns them).
struct ddx_batch {
  bool keyed = false;
  BatchKey key;
  hipStream_t st = nullptr;
  std::vector<BatchMember> members;
  std::vector<ddx_ctx *> done;
  FusedSet fused;
  bool dirty = true;
  int steps = 0;  // batch steps run so far
  // graph replay of the fused step (the first member's ddx_set_graphs):
  // recaptured whenever the members change; d_step is its step counter
  bool graphs = false;
  FusedGraph graph;
  int *d_step = nullptr;
  // RTM members (ddx_batch_add_rtm): those in their forward pass with
  // boundary save and those in their reverse loop, each group with its
  // fused tables (rebuilt when it changes) and the member whose brick
  // shapes the tables use
  struct Rtm {
    ddx_ctx *c;
    int ishot;
    int join;  // time index: forward pass batch step - join; reverse loop join - batch step
  };
  std::vector<Rtm> rf, rr;
  RtmTables tf, tr;
  ddx_ctx *tf_c0 = nullptr, *tr_c0 = nullptr;
  bool dirty_f = true, dirty_r = true;
  int image_fp64 = -1;  // image precision of the RTM members (set by the first)
  std::string err;
};

extern "C" int ddx_batch_create(ddx_batch **out) {
  if (!out) return DDX_EINVAL;
  *out = new ddx_batch();
  return DDX_OK;
}

extern "C" const char *ddx_batch_last_error(const ddx_batch *b) { return b ? b->err.c_str() : ""; }

static bool batch_holds(const ddx_batch *b, const ddx_ctx *c) {
  for (const BatchMember &m : b->members)
    if (m.c == c) return true;
  for (const auto *g : {&b->rf, &b->rr})
    for (const ddx_batch::Rtm &m : *g)
      if (m.c == c) return true;
  return false;
}

extern "C" int ddx_batch_add(ddx_batch *b, ddx_ctx *c, int ishot) {
  if (!b || !c) return DDX_EINVAL;
  if (batch_holds(b, c)) { b->err = "the context is already in the batch"; return DDX_EINVAL; }
  for (ddx_ctx *d : b->done)
    if (d == c) { b->err = "the context's finished shot has not been collected"; return DDX_EINVAL; }
  // the first member to join sets what the others must share
  const bool first = !b->keyed;
  const BatchKey key = first ? batch_key(c) : b->key;
  if (!batch_compatible(key, c)) {
    b->err = kBatchMismatch;
    return DDX_EINVAL;
  }
  hipError_t e = hipSetDevice(key.device);
  if (e == hipSuccess && !b->st) e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
  if (e == hipSuccess && !b->d_step) e = hipMalloc((void **)&b->d_step, sizeof(int));
  if (e != hipSuccess) { b->err = std::string("HIP: ") + hipGetErrorString(e); return DDX_EHIP; }
  BatchMember m;
  if (int rc = batch_prepare(c, ishot, key, m)) {
    b->err = c->err;
    return rc;
  }
  if (first) {
    b->key = key;
    b->keyed = true;
    b->graphs = c->use_graph;
  }
  m.join = b->steps;
  b->members.push_back(std::move(m));
  b->dirty = true;
  return DDX_OK;
}

// Add RTM shot ishot of c (observed data set, as for ddx_rtm_shot): it
// resets both wavefield sets, runs its ntime forward steps with boundary
// save from the next batch step on, then its ntime reverse iterations, and
// is returned by ddx_batch_run when its images hold the shot.
extern "C" int ddx_batch_add_rtm(ddx_batch *b, ddx_ctx *c, int ishot) {
  if (!b || !c) return DDX_EINVAL;
  if (batch_holds(b, c)) { b->err = "the context is already in the batch"; return DDX_EINVAL; }
  for (ddx_ctx *d : b->done)
    if (d == c) { b->err = "the context's finished shot has not been collected"; return DDX_EINVAL; }
  const bool first = !b->keyed;
  const BatchKey key = first ? batch_key(c) : b->key;
  if (!batch_compatible(key, c) || c->set_overlap || (b->image_fp64 >= 0 && c->d.image_fp64 != b->image_fp64)) {
    b->err = std::string(kBatchMismatch) + ", one image precision for the RTM members, and no reverse-set overlap";
    return DDX_EINVAL;
  }
  hipError_t e = hipSetDevice(key.device);
  if (e == hipSuccess && !b->st) e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
  if (e == hipSuccess && !b->d_step) e = hipMalloc((void **)&b->d_step, sizeof(int));
  if (e != hipSuccess) { b->err = std::string("HIP: ") + hipGetErrorString(e); return DDX_EHIP; }
  int rc = rtm_begin(c, ishot);
  if (!rc) rc = ddx_reset_wave(c, 1);
  if (!rc) rc = hipStreamSynchronize(c->st) == hipSuccess ? DDX_OK : DDX_EHIP;
  if (rc) { b->err = c->err; return rc; }
  if (first) {
    b->key = key;
    b->keyed = true;
    b->graphs = c->use_graph;
  }
  b->image_fp64 = c->d.image_fp64;
  b->rf.push_back({c, ishot, b->steps});
  b->dirty_f = true;
  return DDX_OK;
}

// Fused tables of one RTM group (rev 0 forward, 1 reverse); c0 gets the
// member whose brick shapes they use.
static int rtm_group_tables(ddx_batch *b, const std::vector<ddx_batch::Rtm> &g, int rev, RtmTables &T,
                            ddx_ctx *&c0) {
  std::vector<ddx_ctx *> cs;
  std::vector<int> shots, joins;
  for (const ddx_batch::Rtm &m : g) {
    cs.push_back(m.c);
    shots.push_back(m.ishot);
    joins.push_back(m.join);
  }
  c0 = cs[0];
  return rtm_batch_tables(cs, shots, T, b->err, &joins, rev);  // each upload frees the table it replaces
}

// Take a member out before its shot ends (a cancelled or diverged request);
// the context keeps its partial wavefields and seismograms and is the
// caller's again when the call returns.
extern "C" int ddx_batch_remove(ddx_batch *b, ddx_ctx *c) {
  if (!b || !c) return DDX_EINVAL;
  for (size_t q = 0; q < b->members.size(); q++)
    if (b->members[q].c == c) {
      if (b->st) {
        const hipError_t e = hipStreamSynchronize(b->st);
        if (e != hipSuccess) { b->err = std::string("HIP: ") + hipGetErrorString(e); return DDX_EHIP; }
      }
      b->members.erase(b->members.begin() + (long)q);
      b->dirty = true;
      return DDX_OK;
    }
  for (auto *g : {&b->rf, &b->rr})
    for (size_t q = 0; q < g->size(); q++)
      if ((*g)[q].c == c) {
        if (b->st) {
          const hipError_t e = hipStreamSynchronize(b->st);
          if (e != hipSuccess) { b->err = std::string("HIP: ") + hipGetErrorString(e); return DDX_EHIP; }
        }
        g->erase(g->begin() + (long)q);
        b->dirty_f = b->dirty_r = true;
        return DDX_OK;
      }
  b->err = "the context is not a member of the batch";
  return DDX_EINVAL;
}

extern "C" int ddx_batch_size(const ddx_batch *b, int *members) {
  if (!b || !members) return DDX_EINVAL;
  *members = (int)(b->members.size() + b->rf.size() + b->rr.size());
  return DDX_OK;
}

extern "C" int ddx_batch_run(ddx_batch *b, int max_steps, ddx_ctx **done, int cap, int *ndone) {
  if (!b || max_steps < 0 || cap < 0 || !ndone || (cap > 0 && !done)) return DDX_EINVAL;
  *ndone = 0;
  if (b->keyed && hipSetDevice(b->key.device) != hipSuccess) { b->err = "HIP: cannot select the batch's device"; return DDX_EHIP; }
  for (int s = 0; s < max_steps && (!b->members.empty() || !b->rf.empty() || !b->rr.empty()) && b->done.empty();
       s++) {
    // graph replay only while the batch holds modelling shots alone; an
    // eager step drops the captured step (its device step counter is then
    // behind), so the next replayed step recaptures it
    const bool graphs = b->graphs && b->rf.empty() && b->rr.empty();
    if (b->dirty || b->dirty_f || b->dirty_r) {
      hipError_t e = hipStreamSynchronize(b->st);
      if (e != hipSuccess) { b->err = std::string("HIP: ") + hipGetErrorString(e); return DDX_EHIP; }
    }
    if (b->dirty) {
      b->graph.reset();
      if (!b->members.empty()) {
        const hipError_t e = fuse_members(b->members, b->fused);
        if (e != hipSuccess) { b->err = std::string("HIP: ") + hipGetErrorString(e); return DDX_EHIP; }
      }
      b->dirty = false;
    }
    if (!graphs) b->graph.reset();
    if (graphs && !b->members.empty() && !b->graph.x) {
      if (int rc = capture_fused(b->key, b->fused, b->members, b->d_step, b->st, b->graph, b->err)) return rc;
      launch_set_int(b->d_step, b->steps, b->st);
    }
    if (b->dirty_f) {
      if (!b->rf.empty())
        if (int rc = rtm_group_tables(b, b->rf, 0, b->tf, b->tf_c0)) return rc;
      b->dirty_f = false;
    }
    if (b->dirty_r) {
      if (!b->rr.empty())
        if (int rc = rtm_group_tables(b, b->rr, 1, b->tr, b->tr_c0)) return rc;
      b->dirty_r = false;
    }
    if (!b->members.empty()) {
      if (graphs) {
        const hipError_t e = hipGraphLaunch(b->graph.x, b->st);
        if (e != hipSuccess) { b->err = std::string("graph launch failed: ") + hipGetErrorString(e); return DDX_EHIP; }
      } else if (!fused_step(b->key, b->fused, b->members, b->steps, b->st)) {
        b->err = kNoKernel;
        return DDX_EINVAL;
      }
    }
    std::vector<ddx_ctx *> cs;
    std::vector<int> its;
    if (!b->rf.empty()) {
      for (const ddx_batch::Rtm &m : b->rf) { cs.push_back(m.c); its.push_back(b->steps - m.join); }
      if (!rtm_fused_forward(b->key, b->tf_c0, b->tf, cs, its, b->steps, b->st)) { b->err = kNoKernel; return DDX_EINVAL; }
      for (ddx_ctx *c : cs)
        if (int rc = ring_status(c, b->err)) return rc;
    }
    if (!b->rr.empty()) {
      cs.clear();
      its.clear();
      for (const ddx_batch::Rtm &m : b->rr) { cs.push_back(m.c); its.push_back(m.join - b->steps); }
      if (!rtm_fused_reverse(b->key, b->tr_c0, b->tr, cs, its, b->steps, b->st)) { b->err = kNoKernel; return DDX_EINVAL; }
      for (ddx_ctx *c : cs)
        if (int rc = ring_status(c, b->err)) return rc;
    }
    b->steps++;
    // RTM members whose forward pass ended start their reverse loop at the
    // next step (time index ntime - 1); those whose reverse loop ended are done
    {
      std::vector<ddx_batch::Rtm> keep;
      for (ddx_batch::Rtm &m : b->rf) {
        if (b->steps - m.join >= m.c->d.ntime) {
          b->rr.push_back({m.c, m.ishot, b->steps + m.c->d.ntime - 1});
          b->dirty_f = b->dirty_r = true;
        } else {
          keep.push_back(m);
        }
      }
      b->rf.swap(keep);
      keep.clear();
      for (ddx_batch::Rtm &m : b->rr) {
        if (m.join - b->steps < 0) {
          const hipError_t e = hipStreamSynchronize(b->st);
          if (e != hipSuccess) { b->err = std::string("HIP: ") + hipGetErrorString(e); return DDX_EHIP; }
          m.c->cell_updates += (2.0 * ext_cells(m.c) + phy_cells(m.c)) * m.c->d.ntime;
          b->done.push_back(m.c);
          b->dirty_r = true;
        } else {
          keep.push_back(m);
        }
      }
      b->rr.swap(keep);
    }
    bool finished = false;
    for (BatchMember &m : b->members) finished |= ++m.it >= m.c->d.ntime;
    if (finished) {
      const hipError_t e = hipStreamSynchronize(b->st);
      if (e != hipSuccess) { b->err = std::string("HIP: ") + hipGetErrorString(e); return DDX_EHIP; }
      std::vector<BatchMember> keep;
      for (BatchMember &m : b->members) {
        if (m.it >= m.c->d.ntime) {
          m.c->cell_updates += ext_cells(m.c) * m.c->d.ntime;
          b->done.push_back(m.c);
        } else {
          keep.push_back(std::move(m));
        }
      }
      b->members.swap(keep);
      b->dirty = true;
    }
  }
  const hipError_t e = b->st ? hipStreamSynchronize(b->st) : hipSuccess;
  if (e != hipSuccess) { b->err = std::string("HIP: ") + hipGetErrorString(e); return DDX_EHIP; }
  const int k = std::min(cap, (int)b->done.size());
  for (int i = 0; i < k; i++) done[i] = b->done[i];
  b->done.erase(b->done.begin(), b->done.begin() + k);
  *ndone = k;
  return DDX_OK;
}

extern "C" int ddx_batch_destroy(ddx_batch *b) {
  if (!b) return DDX_OK;
  if (b->st) {
    (void)hipStreamSynchronize(b->st);
    b->graph.reset();
    if (b->d_step) (void)hipFree(b->d_step);
    b->fused.v.reset();
    b->fused.s.reset();
    b->fused.p.reset();
    b->fused.io.reset();
    (void)hipStreamDestroy(b->st);
  }
  delete b;
  return DDX_OK;
}

extern "C" int ddx_set_graphs(ddx_ctx *c, int on) {
  if (!c) return DDX_EINVAL;
  c->use_graph = on != 0;
  return DDX_OK;
}

extern "C" int ddx_set_exchange(ddx_ctx *c, int mode) {
  if (!c) return DDX_EINVAL;
  if (mode < 0 || mode > 2) {
    c->err = "exchange mode must be 0 (direct copy), 1 (CMM message blocks) or 2 (pack on store)";
    return DDX_EINVAL;
  }
  if (mode == 2 && !c->have_mirror) {
    c->err = "pack on store needs neighbouring subdomains and the mirror tables (DDX_MIRROR=1 at creation)";
    return DDX_EINVAL;
  }
  if (mode == 1)
    if (int rc = build_cmm(c)) return rc;
  c->exch_mode = mode;
  c->graph_version++;
  return DDX_OK;
}

extern "C" int ddx_get_exchange(const ddx_ctx *c, int *mode) {
  if (!c || !mode) return DDX_EINVAL;
  *mode = c->exch_mode;
  return DDX_OK;
}

extern "C" int ddx_get_seismograms(ddx_ctx *c, int ishot, float *out) {
  if (!c || !out || ishot < 0 || ishot >= c->nshot) return DDX_EINVAL;
  if (!c->seis[ishot]) { c->err = "shot has no seismograms"; return DDX_ESTATE; }
  HIPCHECK(c, hipMemcpyAsync(out, c->seis[ishot], sizeof(float) * 3 * (size_t)c->nrec * c->d.ntime,
                             hipMemcpyDeviceToHost, c->st));
  HIPCHECK(c, hipStreamSynchronize(c->st));
  return DDX_OK;
}

extern "C" int ddx_set_observed(ddx_ctx *c, int ishot, const float *in) {
  if (!c || !in || ishot < 0 || ishot >= c->nshot) return DDX_EINVAL;
  int rc = ensure_seis(c, ishot);
  if (rc) return rc;
  HIPCHECK(c, hipMemcpyAsync(c->seis[ishot], in, sizeof(float) * 3 * (size_t)c->nrec * c->d.ntime,
                             hipMemcpyHostToDevice, c->st));
  HIPCHECK(c, hipStreamSynchronize(c->st));
  return DDX_OK;
}

// The exponent arrays of per-slice ring scaling (once the store exists).
static int ensure_ring_exp(ddx_ctx *c) {
  if (!c->ring_dyn || !ring_scaled(c) || !c->store || c->store_exp) return DDX_OK;
  if (int rc = dalloc(c, (void **)&c->store_exp, sizeof(int) * 10 * (size_t)c->store_slices)) return rc;
  return dalloc(c, (void **)&c->ring_max, sizeof(unsigned) * 10);
}

static int ensure_store(ddx_ctx *c) {
  const long slices = std::max(c->d.ntime - 1, 1);
  if (c->store) return DDX_OK;
  size_t bytes = (size_t)c->ring_esz * (size_t)c->ring_slice * slices;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess && bytes + (256u << 20) > fr) {
    c->err = "VOWR store of " + std::to_string(bytes) + " bytes does not fit in free device memory";
    return DDX_ENOMEM;
  }
  c->store_slices = slices;
  if (int rc = dalloc(c, (void **)&c->store, bytes)) return rc;
  if (c->ring_fmt == STORE_BFP16)
    if (int rc = dalloc(c, (void **)&c->store_bexp, (size_t)c->ring_runs * slices)) return rc;
  return ensure_ring_exp(c);
}

static void ring(ddx_ctx *c, int dir, int which, int it, hipStream_t st) {
  // slices 0 .. store_slices - 1 exist: an index outside them launches
  // nothing and is reported by the caller (ring_oob)
  if (!c->store || it < 0 || it >= c->store_slices) {
    c->ring_oob = true;
    return;
  }
  const DevTable<CopyDesc> &t = c->ring[which];
  char *slice = c->store + (long)it * c->ring_slice * c->ring_esz;
  if (!st) st = c->st;
  if (c->ring_fmt == c->fmt) launch_copy(dir, c->esz, t.d, t.starts, t.n, t.nblk, slice, st);
  else if (c->ring_fmt == STORE_BFP16)
    launch_ring_bfp(dir, t.d, t.starts, t.n, t.nblk, slice, c->store_bexp + (long)it * c->ring_runs, st);
  else if (c->ring_dyn && c->store_exp) {
    int *exps = c->store_exp + (long)it * 10;
    if (dir > 0) {
      launch_ring_max(t.d, t.starts, t.n, t.nblk, c->ring_max, st);
      launch_ring_exp(c->ring_max, exps, 10, st);
    }
    launch_copy_dyn(dir, c->ring_fmt, t.d, t.starts, t.n, t.nblk, slice, exps, st);
  } else
    launch_copy_conv(dir, c->ring_fmt, t.d, t.starts, t.n, t.nblk, slice, std::ldexp(1.0f, c->rlg[0]),
                     std::ldexp(1.0f, c->rlg[1]), st);
}

// DDX_EINVAL (and the error text in err) when a boundary-store copy of c
// asked for a slice outside its store since the last check.
static int ring_status(ddx_ctx *c, std::string &err) {
  if (!c->ring_oob) return DDX_OK;
  c->ring_oob = false;
  err = "internal: a VOWR boundary-store slice outside the store was asked for (no copy made)";
  return DDX_EINVAL;
}

extern "C" int ddx_vowr_save(ddx_ctx *c, int it) {
  if (!c || it < 0 || it > c->d.ntime - 2) return DDX_EINVAL;
  if (int rc = ensure_store(c)) return rc;
  ring(c, +1, 0, it);
  return DDX_OK;
}
extern "C" int ddx_vowr_inject(ddx_ctx *c, int which, int it) {
  if (!c || which < 1 || which > 2 || it < 0 || it > c->d.ntime - 2) return DDX_EINVAL;
  if (!c->store) { c->err = "no VOWR store yet (ddx_vowr_save or ddx_rtm_shot first)"; return DDX_ESTATE; }
  ring(c, -1, which, it);
  return DDX_OK;
}
extern "C" int ddx_recv_inject(ddx_ctx *c, int set, int ishot, int it) {
  if (!c || set < 0 || set > 1 || it < 0 || it >= c->d.ntime) return DDX_EINVAL;
  if (int rc = check_ready(c, ishot)) return rc;
  if (!c->seis[ishot]) { c->err = "no observed data for this shot"; return DDX_ESTATE; }
  launch_recv_inject(c->fmt, c->d_vptr[set], c->d_rsub, c->d_roff, c->nrec, it, c->d.ntime, c->seis[ishot],
                     scale_of(c, 0), c->st);
  return DDX_OK;
}
extern "C" int ddx_image_step(ddx_ctx *c) {
  if (!c) return DDX_EINVAL;
  ProfScope ps(c, 5, c->img.bytes);
  launch_image(c->d.image_fp64, c->fmt, c->img.d, c->img.starts, c->img.n, c->img.nblk, 1.0f / scale_of(c, 0),
               c->st);
  return DDX_OK;
}

// Forward with boundary save (BLEND/...meat.hpp:125-155), then the reverse
// loop (RTM/rtm3d_blend_elastic_model_mpi_meat.hpp:64-113).  The backward
// set reuses the final forward state (BLEND/...meat.hpp:161-170).
static void snapshot(ddx_ctx *c, int pass, int q) {
  const DevTable<CopyDesc> &t = c->snap_tab[c->snap_field];
  char *dst = (char *)c->snap_buf + ((size_t)pass * c->snap_count + q) * c->snap_cells * c->esz;
  launch_copy(+1, c->esz, t.d, t.starts, t.n, t.nblk, dst, c->st);
}

extern "C" int ddx_set_snapshots(ddx_ctx *c, int field, int every) {
  if (!c) return DDX_EINVAL;
  if (every < 0 || (every > 0 && (field < 0 || field >= NF))) {
    c->err = "snapshots: field must be a field id and every >= 0";
    return DDX_EINVAL;
  }
  HIPCHECK(c, hipStreamSynchronize(c->st));
  dfree(c, &c->snap_buf);
  c->snap_every = 0;
  c->snap_count = 0;
  if (every == 0) return DDX_OK;
  if (!c->snap_tab[field].d) {  // gridphy box of every subdomain, rank after rank
    std::vector<CopyDesc> h;
    std::vector<int> st;
    int blk = 0;
    long off = 0;
    for (int s = 0; s < c->nsub; s++) {
      const Geom &g = c->g[s];
      CopyDesc p;
      p.field = fptr(c, 0, s, field);
      p.sy = g.sy; p.sz = g.sz;
      p.i0 = g.mlo[0]; p.j0 = g.mlo[1]; p.k0 = g.mlo[2];
      p.ni = g.pn[0]; p.nj = g.pn[1]; p.nk = g.pn[2];
      p.buf_off = off;
      p.bi = p.bj = p.bk = 0;
      p.bni = p.ni; p.bnj = p.nj;
      p.cls = field <= VZ ? 0 : 1;
      p.fid = 0;
      p.blk0 = blk;
      st.push_back(blk);
      const long cnt = (long)g.pn[0] * g.pn[1] * g.pn[2];
      blk += (int)((cnt + kBlock - 1) / kBlock);
      off += cnt;
      h.push_back(p);
    }
    int rc = upload_table(c, c->snap_tab[field], h, st, blk);
    if (rc) return rc;
    c->snap_cells = off;
  }
  const int count = (c->d.ntime - 1) / every + 1;
  int rc = dalloc(c, (void **)&c->snap_buf, (size_t)c->esz * 2 * (size_t)count * c->snap_cells);
  if (rc) return rc;
  c->snap_field = field;
  c->snap_every = every;
  c->snap_count = count;
  return DDX_OK;
}

extern "C" int ddx_get_snapshot(ddx_ctx *c, int pass, int index, float *out) {
  if (!c || !out || pass < 0 || pass > 1) return DDX_EINVAL;
  if (c->snap_every <= 0 || index < 0 || index >= c->snap_count) {
    c->err = "no such snapshot (ddx_set_snapshots first; index < (ntime - 1) / every + 1)";
    return DDX_EINVAL;
  }
  std::vector<char> raw((size_t)c->snap_cells * c->esz);
  const char *src = (const char *)c->snap_buf + ((size_t)pass * c->snap_count + index) * c->snap_cells * c->esz;
  HIPCHECK(c, hipMemcpyAsync(raw.data(), src, raw.size(), hipMemcpyDeviceToHost, c->st));
  HIPCHECK(c, hipStreamSynchronize(c->st));
  const int cls = field_class(c->snap_field);
  const float inv = cls >= 0 ? 1.0f / scale_of(c, cls) : 1.0f;
  std::vector<float> h((size_t)c->snap_cells);
  for (size_t i = 0; i < h.size(); i++) h[i] = load_elem(c, raw.data(), (long)i, inv);
  // scatter the per-subdomain blocks into the global gridphy (x fastest)
  size_t off = 0;
  for (int s = 0; s < c->nsub; s++) {
    const Geom &g = c->g[s];
    for (int k = 0; k < g.pn[2]; k++)
      for (int j = 0; j < g.pn[1]; j++) {
        const size_t dst = (size_t)g.poff[0] +
                           (size_t)c->d.nx * ((size_t)(g.poff[1] + j) + (size_t)c->d.ny * (size_t)(g.poff[2] + k));
        memcpy(out + dst, h.data() + off, sizeof(float) * g.pn[0]);
        off += g.pn[0];
      }
  }
  return DDX_OK;
}

// One forward step with the boundary save (BLEND/...meat.hpp:125-155).
static int rtm_forward_step(ddx_ctx *c, int ishot, int it) {
  const int nt = c->d.ntime;
  int rc;
  if ((rc = vel_half(c, 0, RTM_FORWARD))) return rc;
  if ((rc = ddx_source(c, 0, ishot, it, +1))) return rc;
  if ((rc = stress_half(c, 0, RTM_FORWARD))) return rc;
  if (it <= nt - 2) ring(c, +1, 0, it);
  // forward state after step it == backward state after reverse iteration it + 1
  if (c->snap_every > 0 && (nt - 1 - it) % c->snap_every == 0) snapshot(c, 0, (nt - 1 - it) / c->snap_every);
  return DDX_OK;
}

// One reverse iteration it (RTM/rtm3d_blend_elastic_model_mpi_meat.hpp:64-113).
static int rtm_reverse_step(ddx_ctx *c, int ishot, int it) {
  const int nt = c->d.ntime;
  int rc;
  const hipStream_t main_st = c->st;
  if (c->set_overlap) {
    // receiver side (set 1) on its own stream from here to the imaging;
    // it only depends on the previous iteration's imaging
    HIPCHECK(c, hipEventRecord(c->ev_rev_fork, main_st));
    HIPCHECK(c, hipStreamWaitEvent(c->st_rev, c->ev_rev_fork, 0));
    c->st = c->st_rev;
    launch_recv_inject(c->fmt, c->d_vptr[1], c->d_rsub, c->d_roff, c->nrec, it, nt, c->seis[ishot], scale_of(c, 0),
                       c->st);
    rc = vel_half(c, 1, RTM_REVERSE);
    if (!rc) rc = stress_half(c, 1, RTM_REVERSE);
    if (!rc && hipEventRecord(c->ev_rev_join, c->st) != hipSuccess) rc = DDX_EHIP;
    c->st = main_st;
    if (rc) return rc;
    // source side (set 0) on the main stream
    if ((rc = ddx_source(c, 0, ishot, it, -1))) return rc;
    if ((rc = stress_update(c, 0, RTM_BACKWARD))) return rc;
    if (it - 1 >= 0) ring(c, -1, 2, it - 1);
    exchange_after(c, 0, 2, c->tiling_stress[RTM_BACKWARD], c->st);
    if ((rc = vel_update(c, 0, RTM_BACKWARD))) return rc;
    if (it - 1 >= 0) ring(c, -1, 1, it - 1);
    exchange_after(c, 0, 0, c->tiling_vel[RTM_BACKWARD], c->st);
    HIPCHECK(c, hipStreamWaitEvent(main_st, c->ev_rev_join, 0));
  } else {
    if ((rc = ddx_source(c, 0, ishot, it, -1))) return rc;
    launch_recv_inject(c->fmt, c->d_vptr[1], c->d_rsub, c->d_roff, c->nrec, it, nt, c->seis[ishot], scale_of(c, 0),
                       c->st);
    // backward stress, ring inject of step it-1, then the halo exchange
    if ((rc = stress_update(c, 0, RTM_BACKWARD))) return rc;
    if (it - 1 >= 0) ring(c, -1, 2, it - 1);
    exchange_after(c, 0, 2, c->tiling_stress[RTM_BACKWARD], c->st);
    if ((rc = vel_half(c, 1, RTM_REVERSE))) return rc;
    if ((rc = vel_update(c, 0, RTM_BACKWARD))) return rc;
    if (it - 1 >= 0) ring(c, -1, 1, it - 1);
    exchange_after(c, 0, 0, c->tiling_vel[RTM_BACKWARD], c->st);
    if ((rc = stress_half(c, 1, RTM_REVERSE))) return rc;
  }
  {
    ProfScope ps(c, 5, c->img.bytes);
    launch_image(c->d.image_fp64, c->fmt, c->img.d, c->img.starts, c->img.n, c->img.nblk, 1.0f / scale_of(c, 0),
                 c->st);
  }
  if (c->snap_every > 0 && it >= 1 && (nt - it) % c->snap_every == 0) snapshot(c, 1, (nt - it) / c->snap_every);
  return DDX_OK;
}

static int rtm_begin(ddx_ctx *c, int ishot) {
  int rc = check_ready(c, ishot);
  if (rc) return rc;
  if (!c->seis[ishot]) { c->err = "no observed data for this shot"; return DDX_ESTATE; }
  if ((rc = ensure_store(c))) return rc;
  return ddx_reset_wave(c, 0);
}

extern "C" int ddx_rtm_shot(ddx_ctx *c, int ishot) {
  if (!c) return DDX_EINVAL;
  int rc = rtm_begin(c, ishot);
  if (rc) return rc;
  const int nt = c->d.ntime;
  for (int it = 0; it < nt; it++)
    if ((rc = rtm_forward_step(c, ishot, it))) return rc;
  if ((rc = ddx_reset_wave(c, 1))) return rc;
  for (int it = nt - 1; it >= 0; it--)
    if ((rc = rtm_reverse_step(c, ishot, it))) return rc;
  if ((rc = ring_status(c, c->err))) return rc;
  HIPCHECK(c, hipGetLastError());
  c->cell_updates += (2.0 * ext_cells(c) + phy_cells(c)) * nt;
  return DDX_OK;
}

// Chunked shots (no reference counterpart): the steps of one modelling shot
// (ntime stages) or one RTM shot (2 ntime stages: the forward steps, then
// the reverse iterations ntime-1 .. 0) run in pieces, stage s0 onward; a
// piece that starts at 0 starts the shot, every other piece continues it.
// Between pieces the caller may check the wavefields (ddx_check_finite) or
// stop.  The pieces together run exactly the whole-shot call's kernels.
static int chunk_order(ddx_ctx *c, int kind, int ishot, int s0, int n, int stages) {
  if (s0 < 0 || n < 0 || s0 + n > stages) { c->err = "stages out of range"; return DDX_EINVAL; }
  if (s0 > 0 && (c->chunk_kind != kind || c->chunk_shot != ishot || c->chunk_next != s0)) {
    c->err = "a piece must continue the shot where the previous piece stopped";
    return DDX_ESTATE;
  }
  return DDX_OK;
}

extern "C" int ddx_model_steps(ddx_ctx *c, int ishot, int s0, int n) {
  if (!c) return DDX_EINVAL;
  int rc = check_ready(c, ishot);
  if (rc) return rc;
  const int nt = c->d.ntime;
  if ((rc = chunk_order(c, 1, ishot, s0, n, nt))) return rc;
  if (s0 == 0) {
    if ((rc = ensure_seis(c, ishot))) return rc;
    if ((rc = ddx_reset_wave(c, 0))) return rc;
  }
  c->chunk_kind = 1;
  c->chunk_shot = ishot;
  for (int it = s0; it < s0 + n; it++) {
    if ((rc = model_step(c, ishot, it, nullptr))) { c->chunk_kind = 0; return rc; }
    c->chunk_next = it + 1;
  }
  HIPCHECK(c, hipGetLastError());
  if (c->chunk_next == nt && n > 0) c->cell_updates += ext_cells(c) * nt;
  return DDX_OK;
}

extern "C" int ddx_rtm_steps(ddx_ctx *c, int ishot, int s0, int n) {
  if (!c) return DDX_EINVAL;
  const int nt = c->d.ntime;
  int rc;
  if ((rc = chunk_order(c, 2, ishot, s0, n, 2 * nt))) return rc;
  if (s0 == 0 && (rc = rtm_begin(c, ishot))) return rc;
  c->chunk_kind = 2;
  c->chunk_shot = ishot;
  for (int st = s0; st < s0 + n; st++) {
    if (st < nt) rc = rtm_forward_step(c, ishot, st);
    else {
      if (st == nt) rc = ddx_reset_wave(c, 1);
      if (!rc) rc = rtm_reverse_step(c, ishot, 2 * nt - 1 - st);
    }
    if (rc) { c->chunk_kind = 0; return rc; }
    c->chunk_next = st + 1;
  }
  if ((rc = ring_status(c, c->err))) { c->chunk_kind = 0; return rc; }
  HIPCHECK(c, hipGetLastError());
  if (c->chunk_next == 2 * nt && n > 0) c->cell_updates += (2.0 * ext_cells(c) + phy_cells(c)) * nt;
  return DDX_OK;
}

// Any non-finite value in the velocities of either wavefield set (a run
// that diverged: an unstable time step, or a model outside the CFL limit
// with cfl_policy = 1).  One pass over 6 fields of every subdomain.
extern "C" int ddx_check_finite(ddx_ctx *c, int *finite) {
  if (!c || !finite) return DDX_EINVAL;
  HIPCHECK(c, hipMemsetAsync(c->d_flag, 0, sizeof(int), c->st));
  for (int set = 0; set < 2; set++)
    for (int s = 0; s < c->nsub; s++)
      for (int f = VX; f <= VZ; f++)
        launch_nonfinite(c->fmt, c->wave[set][s] + (long)f * c->g[s].fpitch * c->esz, c->g[s].nstore, c->d_flag,
                         c->st);
  int h = 0;
  HIPCHECK(c, hipMemcpyAsync(&h, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->st));
  HIPCHECK(c, hipStreamSynchronize(c->st));
  *finite = h ? 0 : 1;
  return DDX_OK;
}

extern "C" int ddx_finalize_images(ddx_ctx *c) {
  if (!c) return DDX_EINVAL;
  launch_image_post(c->d.image_fp64, c->img.d, c->img.starts, c->img.n, c->img.nblk, c->st);
  HIPCHECK(c, hipGetLastError());
  HIPCHECK(c, hipStreamSynchronize(c->st));
  return DDX_OK;
}

extern "C" int ddx_reset_images(ddx_ctx *c) {
  if (!c) return DDX_EINVAL;
  const size_t esz = c->d.image_fp64 ? sizeof(double) : sizeof(float);
  for (int s = 0; s < c->nsub; s++) {
    const Geom &g = c->g[s];
    const size_t np = (size_t)g.pn[0] * g.pn[1] * g.pn[2];
    for (int q = 0; q < 9; q++) HIPCHECK(c, hipMemsetAsync(c->images[(size_t)s * 9 + q], 0, esz * np, c->st));
  }
  return DDX_OK;
}

extern "C" int ddx_get_image(ddx_ctx *c, int which, int rank, float *out) {
  if (!c || !out || which < 0 || which >= DDX_NIMAGE || rank < 0 || rank >= c->nsub) return DDX_EINVAL;
  const Geom &g = c->g[rank];
  const size_t np = (size_t)g.pn[0] * g.pn[1] * g.pn[2];
  void *src = c->images[(size_t)rank * 9 + which];
  if (c->d.image_fp64) {
    std::vector<double> h(np);
    HIPCHECK(c, hipMemcpyAsync(h.data(), src, sizeof(double) * np, hipMemcpyDeviceToHost, c->st));
    HIPCHECK(c, hipStreamSynchronize(c->st));
    for (size_t i = 0; i < np; i++) out[i] = (float)h[i];
  } else {
    HIPCHECK(c, hipMemcpyAsync(out, src, sizeof(float) * np, hipMemcpyDeviceToHost, c->st));
    HIPCHECK(c, hipStreamSynchronize(c->st));
  }
  return DDX_OK;
}

// Assemble / scatter a field on the global extended grid.
static int field_io(ddx_ctx *c, int set, int field, float *out, const float *in) {
  if (!c || set < 0 || set > 1 || field < 0 || field >= NF) return DDX_EINVAL;
  const int cls = field_class(field);
  const float sc = cls >= 0 ? scale_of(c, cls) : 1.0f, inv = 1.0f / sc;
  const long GX = c->gext[0], GY = c->gext[1];
  if (out) std::fill(out, out + GX * GY * (long)c->gext[2], 0.f);
  for (int s = 0; s < c->nsub; s++) {
    const Geom &g = c->g[s];
    std::vector<char> h((size_t)g.nstore * c->esz);
    char *base = c->wave[set][s] + (long)field * g.fpitch * c->esz;
    HIPCHECK(c, hipMemcpyAsync(h.data(), base, h.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHECK(c, hipStreamSynchronize(c->st));
    for (int k = 0; k < g.en[2]; k++)
      for (int j = 0; j < g.en[1]; j++)
        for (int i = 0; i < g.en[0]; i++) {
          const long o = g.origin + i + g.sy * j + g.sz * k;
          const long q = (i + g.ext_goff[0]) + GX * ((j + g.ext_goff[1]) + GY * (k + g.ext_goff[2]));
          if (out) out[q] = load_elem(c, h.data(), o, inv);
          else store_elem(c, h.data(), o, in[q], sc);
        }
    if (in) {
      HIPCHECK(c, hipMemcpyAsync(base, h.data(), h.size(), hipMemcpyHostToDevice, c->st));
      HIPCHECK(c, hipStreamSynchronize(c->st));
    }
  }
  // ghosts follow the owners
  if (in) {
    const int batch = field <= VZ ? 0 : (field == SAU ? 2 : 1);
    exchange(c, set, batch);
    HIPCHECK(c, hipStreamSynchronize(c->st));
  }
  return DDX_OK;
}

extern "C" int ddx_get_field(ddx_ctx *c, int set, int field, float *out) {
  if (!out) return DDX_EINVAL;
  return field_io(c, set, field, out, nullptr);
}
extern "C" int ddx_set_field(ddx_ctx *c, int set, int field, const float *in) {
  if (!in) return DDX_EINVAL;
  return field_io(c, set, field, nullptr, in);
}

extern "C" int ddx_synchronize(ddx_ctx *c) {
  if (!c) return DDX_EINVAL;
  HIPCHECK(c, hipStreamSynchronize(c->st));
  HIPCHECK(c, hipGetLastError());
  return DDX_OK;
}

extern "C" int ddx_set_tiling(ddx_ctx *c, int tx, int ty, int cy, int kz) {
  if (!c) return DDX_EINVAL;
  if (c->fmt != STORE_F32 && !(kz > 0 && cy > 0)) {
    c->err = "the 16-bit storage formats run the z-blocked bricks only";
    return DDX_EINVAL;
  }
  for (int k = 0; k < 2; k++)
    for (int m = 0; m < kModes; m++) {
      if (m == FWI && !(kz > 0 && cy > 0)) continue;  // FWI keeps a z-blocked brick
      int rc = ddx_set_tiling_mode(c, k, m, tx, ty, cy, kz);
      if (rc) return rc;
    }
  return DDX_OK;
}

extern "C" int ddx_set_tiling_mode(ddx_ctx *c, int kernel, int mode, int tx, int ty, int cy, int kz) {
  if (c) c->graph_version++;
  if (!c) return DDX_EINVAL;
  Tiling t{tx, ty, cy, kz};
  if (kernel < 0 || kernel > 1 || mode < 0 || mode >= kModes || (mode == FWI && !(kz > 0 && cy > 0))) {
    c->err = "kernel must be 0 (velocity) or 1 (stress) and mode 0..3";
    return DDX_EINVAL;
  }
  if (kz < 0 || (kz == 0 && !(tx == kTileX && ty == kTileY && cy == 1 && c->fmt == STORE_F32)) ||
      (kz > 0 && !tiling_supported(t, c->fmt))) {
    c->err = "unsupported tile shape";
    return DDX_EINVAL;
  }
  HIPCHECK(c, hipStreamSynchronize(c->st));
  (kernel == 0 ? c->tiling_vel : c->tiling_stress)[mode] = t;
  for (int m = 0; m < 2; m++)
    if (!c->model[m].empty() && !phy_tables(c, t, m)) return DDX_EHIP;
  return DDX_OK;
}

// Brick autotuner (no counterpart in the reference, which launches one
// thread per cell): for each physical-region kernel (velocity, stress) and
// propagation mode, time `reps` launches of every z-blocked shape of the
// context's storage format (one-cell and paired bricks) with HIP events on
// the context stream, and keep the fastest.  Results are cached per process
// by storage format, decomposition and subdomain size bucket (each gridphy
// axis rounded up to 32 cells), so contexts of similar size reuse them.  The
// timing launches run on the current wavefields; both sets are reset
// afterwards.
static std::mutex g_tune_mu;
static std::map<std::string, Tiling> g_tune_cache;

static std::string tune_key(const ddx_ctx *c, int kernel, int mode) {
  char b[160];
  int big[3] = {0, 0, 0};
  for (const Geom &g : c->g)
    for (int a = 0; a < 3; a++) big[a] = std::max(big[a], (g.pn[a] + 31) / 32 * 32);
  snprintf(b, sizeof b, "f%d k%d m%d n%d R%d b%dx%dx%d mig%d", c->fmt, kernel, mode, c->nsub,
           mode == RTM_BACKWARD ? c->d.rec_width : 0, big[0], big[1], big[2], (int)model_for(c, mode));
  return b;
}

extern "C" int ddx_autotune(ddx_ctx *c, int reps, int *timed) {
  if (c) c->graph_version++;
  if (!c || reps < 1 || reps > 100) return DDX_EINVAL;
  if (!c->have_model) { c->err = "set the model before ddx_autotune"; return DDX_ESTATE; }
  HIPCHECK(c, hipStreamSynchronize(c->st));
  Tiling cand[64];
  const int nc = std::min(tiling_list(c->fmt, cand, 64), 64);
  hipEvent_t e0, e1;
  HIPCHECK(c, hipEventCreate(&e0));
  HIPCHECK(c, hipEventCreate(&e1));
  int count = 0, rc = DDX_OK;
  const int prof_mask = c->prof.mask;
  c->prof.mask = 0;  // keep the tuning launches out of the measurement
  for (int kernel = 0; kernel < 2 && !rc; kernel++)
    for (int mode = 0; mode < kModes && !rc; mode++) {
      if (mode == FWI) continue;  // FWI keeps its one-cell brick
      Tiling &slot = (kernel == 0 ? c->tiling_vel : c->tiling_stress)[mode];
      const std::string key = tune_key(c, kernel, mode);
      {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(key);
        if (it != g_tune_cache.end()) {
          slot = it->second;
          continue;
        }
      }
      float best = 1e30f;
      Tiling bt = slot;
      for (int q = 0; q < nc && !rc; q++) {
        if ((rc = phy_launch(c, kernel, 0, mode, cand[q]))) break;  // warm-up (and table build)
        if (hipEventRecord(e0, c->st) != hipSuccess) { rc = DDX_EHIP; break; }
        for (int r = 0; r < reps && !rc; r++) rc = phy_launch(c, kernel, 0, mode, cand[q]);
        if (rc || hipEventRecord(e1, c->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
          rc = rc ? rc : DDX_EHIP;
          break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        count++;
        if (ms < best) { best = ms; bt = cand[q]; }
      }
      if (rc) break;
      slot = bt;
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tune_cache[key] = bt;
    }
  c->prof.mask = prof_mask;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    if (c->err.empty()) c->err = "autotune launch failed";
    return rc;
  }
  if (timed) *timed = count;
  if ((rc = ddx_reset_wave(c, 0)) || (rc = ddx_reset_wave(c, 1))) return rc;
  HIPCHECK(c, hipStreamSynchronize(c->st));
  HIPCHECK(c, hipGetLastError());
  return DDX_OK;
}

// Brick shape of one kernel and mode including its lane form (1: one cell
// per thread, 2: two x cells per thread, 3: one cell per thread with the x
// arm from neighbouring lanes), and the shape in use (out[5] = tx, ty, cy,
// kz, form).
extern "C" int ddx_set_brick(ddx_ctx *c, int kernel, int mode, int tx, int ty, int cy, int kz, int form) {
  if (c) c->graph_version++;
  if (!c) return DDX_EINVAL;
  if (form == 1) return ddx_set_tiling_mode(c, kernel, mode, tx, ty, cy, kz);
  Tiling t{tx, ty, cy, kz};
  t.px = form == 2 ? 2 : (form == 3 || form == 4) ? 1 : 0;
  t.xl = form == 3 ? 1 : form == 4 ? 2 : 0;
  if (kernel < 0 || kernel > 1 || mode < 0 || mode >= kModes || mode == FWI || !tiling_supported(t, c->fmt)) {
    c->err = "unsupported brick (forms 2 and 3: see the supported shapes; not the FWI mode)";
    return DDX_EINVAL;
  }
  HIPCHECK(c, hipStreamSynchronize(c->st));
  (kernel == 0 ? c->tiling_vel : c->tiling_stress)[mode] = t;
  for (int m = 0; m < 2; m++)
    if (!c->model[m].empty() && !phy_tables(c, t, m)) return DDX_EHIP;
  return DDX_OK;
}

extern "C" int ddx_get_brick(const ddx_ctx *c, int kernel, int mode, int *out5) {
  if (!c || !out5 || kernel < 0 || kernel > 1 || mode < 0 || mode >= kModes) return DDX_EINVAL;
  const Tiling &t = (kernel == 0 ? c->tiling_vel : c->tiling_stress)[mode];
  out5[0] = t.tx; out5[1] = t.ty; out5[2] = t.cy; out5[3] = t.kz; out5[4] = t.xl == 2 ? 4 : t.xl ? 3 : t.px;
  return DDX_OK;
}

extern "C" int ddx_profile(ddx_ctx *c, int kernel) {
  if (!c) return DDX_EINVAL;
  (void)hipStreamSynchronize(c->st);
  c->prof.cls.clear();
  for (double &b : c->prof.bytes) b = 0;
  c->prof.kernel = kernel >= 0 && kernel < kClasses ? kernel : kernel == -1 ? -1 : -2;
  c->prof.mask = kernel == -1 ? (1 << kClasses) - 1 : (kernel >= 0 && kernel < kClasses ? 1 << kernel : 0);
  return DDX_OK;
}

// Summed milliseconds and launches of the classes in `mask` since ddx_profile.
static int class_time(ddx_ctx *c, int mask, double *ms, long long *launches, double *bytes) {
  HIPCHECK(c, hipStreamSynchronize(c->st));
  double tot = 0, by = 0;
  long long n = 0;
  for (size_t i = 0; i < c->prof.cls.size(); i++) {
    if (!((mask >> c->prof.cls[i]) & 1)) continue;
    float t = 0;
    (void)hipEventElapsedTime(&t, c->prof.pool[i].first, c->prof.pool[i].second);
    tot += t;
    n++;
  }
  for (int k = 0; k < kClasses; k++)
    if ((mask >> k) & 1) by += c->prof.bytes[k];
  if (ms) *ms = tot;
  if (launches) *launches = n;
  if (bytes) *bytes = by;
  return DDX_OK;
}

extern "C" int ddx_kernel_time(ddx_ctx *c, double *ms, long long *launches, double *bytes) {
  if (!c) return DDX_EINVAL;
  return class_time(c, c->prof.mask, ms, launches, bytes);
}

extern "C" int ddx_kernel_class_time(ddx_ctx *c, int kernel, double *ms, long long *launches, double *bytes) {
  if (!c || kernel < 0 || kernel >= kClasses) return DDX_EINVAL;
  return class_time(c, 1 << kernel, ms, launches, bytes);
}

extern "C" int ddx_cell_updates(const ddx_ctx *c, double *cells) {
  if (!c || !cells) return DDX_EINVAL;
  *cells = c->cell_updates;
  return DDX_OK;
}

extern "C" int ddx_device_bytes(const ddx_ctx *c, double *bytes) {
  if (!c || !bytes) return DDX_EINVAL;
  *bytes = c->dev_bytes;
  return DDX_OK;
}

// ---------------------------------------------------------------------------
// host-only helpers
// ---------------------------------------------------------------------------
// Box smoothing of a model on its grid, along x, then y, then z: the mean
// over [i - r, i + r] clipped to the grid, accumulated in double.  This is
// the smoothing BASE/fdtd3d_base_3_geomodel_meat.hpp:7-80 applies to build a
// migration background (smooth_model_x_dir; its y pass reads a fixed offset
// instead of the window and is applied in place, and the z radius is never
// passed, so here all three passes are the x pass along their own axis).
extern "C" int ddx_smooth_model(const float *in, float *out, int nx, int ny, int nz, int rx, int ry, int rz) {
  if (!in || !out || nx <= 0 || ny <= 0 || nz <= 0 || rx < 0 || ry < 0 || rz < 0) return DDX_EINVAL;
  const size_t n = (size_t)nx * ny * nz;
  std::vector<float> a(in, in + n), b(n);
  const int len[3] = {nx, ny, nz}, rad[3] = {rx, ry, rz};
  const size_t stride[3] = {1, (size_t)nx, (size_t)nx * ny};
  for (int ax = 0; ax < 3; ax++) {
    const int L = len[ax], r = rad[ax];
    const size_t st = stride[ax];
    if (r == 0) continue;
    for (size_t base = 0; base < n; base++) {
      const int pos = (int)((base / st) % (size_t)L);
      if (pos != 0) continue;  // walk each line from its first cell
      for (int p = 0; p < L; p++) {
        const int lo = std::max(p - r, 0), hi = std::min(p + r, L - 1);
        double sum = 0.0;
        for (int q = lo; q <= hi; q++) sum += a[base + (size_t)q * st];
        b[base + (size_t)p * st] = (float)(sum / (hi - lo + 1));
      }
    }
    std::swap(a, b);
  }
  std::copy(a.begin(), a.end(), out);
  return DDX_OK;
}

// Device bytes a context will hold for this description (no GPU needed):
// both wavefield sets, the model (twice with a migration model), the CPML
// memory variables and coefficients, the images, the seismograms of nshot
// shots with nrec receivers, and for RTM the VOWR boundary store.  Tables,
// receiver lists and the wavelet (under 1 MB) are left out.
extern "C" int ddx_estimate_bytes(const ddx_desc *d, int nrec, int nshot, int rtm, int migration, double *bytes) {
  if (!d || !bytes || nrec < 0 || nshot < 0) return DDX_EINVAL;
  const int nsub = d->x_divide_num * d->y_divide_num * d->z_divide_num;
  if (nsub <= 0 || d->ntime <= 0) return DDX_EINVAL;
  const double esz = d->image_fp64 ? 8.0 : 4.0;
  const double fsz = d->storage == DDX_STORE_FP32 ? 4.0 : 2.0;  // wavefield element
  const int models = migration ? 2 : 1;
  double b = 0.0;
  for (int r = 0; r < nsub; r++) {
    Geom g;
    std::string err;
    if (!make_geom(*d, r, g, err)) return DDX_EINVAL;
    b += (2.0 * NF * fsz + 4.0 * 3.0 * models) * (double)g.fpitch;
    b += 9.0 * esz * (double)g.pn[0] * g.pn[1] * g.pn[2];
    for_each_slab(g, [&](int i0, int i1, int j0, int j1, int k0, int k1, int axes) {
      if (i1 <= i0 || j1 <= j0 || k1 <= k0) return;
      const double ncell = (double)(i1 - i0) * (j1 - j0) * (k1 - k0);
      b += ncell * __builtin_popcount(axes) * (4.0 * 4.0 * models + 12.0 * fsz);  // coefficients + 2 sets of psi
    });
    if (rtm) {
      long ring = 0;
      for (int fi = 0; fi < 10; fi++)
        for (int side = 0; side < 6; side++) {
          const int a = kSideAxis[side], hi = kSideHi[side];
          if ((hi ? g.nhi[a] : g.nlo[a]) >= 0) continue;
          const int w = ring_w(kRingFields[fi], side, d->rec_width);
          if (w <= 0) continue;
          long n[3] = {g.pn[0], g.pn[1], g.pn[2]};
          n[a] = w;
          const long cnt = n[0] * n[1] * n[2];
          ring += d->ring_storage == DDX_STORE_BFP16 ? (cnt + kRingBlock - 1) / kRingBlock * kRingBlock : cnt;
        }
      // 16-bit entries, plus one exponent byte per run of the block-scaled store
      const double rsz = d->ring_storage == DDX_STORE_FP32 ? fsz
                         : d->ring_storage == DDX_STORE_BFP16 ? 2.0 + 1.0 / kRingBlock : 2.0;
      b += rsz * (double)ring * std::max(d->ntime - 1, 1);
    }
  }
  b += 4.0 * 3.0 * nrec * (double)d->ntime * nshot;
  *bytes = b;
  return DDX_OK;
}

extern "C" int ddx_subdomain(const ddx_desc *desc, int rank, int *out) {
  if (!desc || !out) return DDX_EINVAL;
  Geom g;
  std::string e;
  if (!make_geom(*desc, rank, g, e)) return DDX_EINVAL;
  for (int a = 0; a < 3; a++) {
    out[a] = g.pn[a];
    out[3 + a] = g.poff[a];
    out[6 + a] = g.en[a];
  }
  for (int side = 0; side < 6; side++) {
    const int a = kSideAxis[side], hi = kSideHi[side];
    out[9 + side] = hi ? g.mhi[a] : g.mlo[a];
    out[15 + side] = hi ? g.nhi[a] : g.nlo[a];
  }
  return DDX_OK;
}

extern "C" int ddx_fd_coef(int half, float *out) {
  if (half < 1 || half > 8 || !out) return DDX_EINVAL;
  fd_coef(half, out);
  return DDX_OK;
}

extern "C" int ddx_ntime(float T, float dt) { return (int)std::ceil(T / dt); }

// BASE/fdtd3d_base_4_seisrecord_meat.hpp:273-285 (Ricker, double pow/exp on
// float arguments).
extern "C" int ddx_ricker(float fm, float dt, int ntime, float *out) {
  if (!out || ntime < 0 || fm <= 0 || dt <= 0) return DDX_EINVAL;
  const float dura = 1.3f / fm;
  const int n0 = (int)std::floor(dura / dt + 0.5f);
  for (int i = 0; i < ntime; i++) {
    const double x = kPi * (double)fm * (double)((float)(i - n0) * dt);
    const double p = std::pow(x, 2);
    out[i] = (float)(1.0e4f * (1.0f - 2.0f * p) * std::exp((-1) * p));
  }
  return DDX_OK;
}

// PHY/fdtd3d_phy_mpi_1_host_meat.hpp:142-147: dt_cal = 1 / sum|c| / (s vmax).
extern "C" float ddx_stable_dt(float vmax, float dx, float dy, float dz) {
  float c[4];
  fd_coef(4, c);
  float asum = 0.f;
  for (int i = 0; i < 4; i++) asum += std::fabs(c[i]);
  const float s = std::sqrt(1.0f / (dx * dx) + 1.0f / (dy * dy) + 1.0f / (dz * dz));
  return 1.0f / asum / (s * vmax);
}
