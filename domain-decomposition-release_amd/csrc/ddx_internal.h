// Internal device-side descriptors shared by the kernels (kernels.hip) and the
// host engine (engine.cpp).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace ddx {

constexpr int NF = 19;
// the last six are the FWI accumulators (displacement, dt x velocity increment)
enum : int { VX = 0, VY, VZ, SXX, SYY, SZZ, SXY, SXZ, SYZ, SAU, VXP, VYP, VZP, UX, UY, UZ, VXRT, VYRT, VZRT };
enum : int { PURE_FORWARD = 0, RTM_FORWARD = 1, RTM_BACKWARD = 2, RTM_REVERSE = 3, FWI = 4 };
constexpr int kModes = 5;
// Wavefield storage formats (fields, CPML memory variables, VOWR store): FP32,
// or 16-bit codes of the value times a power-of-two scale per field class.
enum : int { STORE_F32 = 0, STORE_F16 = 1, STORE_BF16 = 2, STORE_BFP16 = 3 };
// STORE_BFP16 (VOWR store only): 16-bit codes (sign, 3-bit exponent, 12-bit
// mantissa) relative to one exponent shared by kRingBlock consecutive
// entries of a slice; each ring box starts on a run.
constexpr int kRingBlock = 16;
// the RTM modes update sau in the physical region
__host__ __device__ constexpr bool rtm_mode(int m) { return m == RTM_FORWARD || m == RTM_BACKWARD || m == RTM_REVERSE; }

// Threads per block of every grid-mapped kernel; the physical-region kernels
// use a 64 (x) by 4 (y) tile so one wavefront covers one contiguous x row.
constexpr int kBlock = 256;
constexpr int kTileX = 64;
constexpr int kTileY = 4;
// Work split of the CPML kernels (a context's pml_cz):
//   1, 2, 4, 8: kBlock consecutive cells of a slab plane per block, cz
//     planes per thread, the x and y stencil taps read through the caches;
//   kPmlTiled + p (p = 2, 4, 8): a tile of pml_tile_x(ni) x (kBlock /
//     pml_tile_x(ni)) plane cells per block, staged in LDS with its halo,
//     p planes marched per block (FP32 storage only).
constexpr int kPmlTiled = 16;
__host__ __device__ constexpr bool pml_tiled(int cz) { return cz > kPmlTiled; }
__host__ __device__ constexpr int pml_planes(int cz) { return cz > kPmlTiled ? cz - kPmlTiled : cz; }
// Tile width: the whole slab row when it is at most 64 cells (the 18-cell
// x slabs), else 64 (one wavefront per row).
__host__ __device__ constexpr int pml_tile_x(int ni) { return ni < 4 ? 4 : ni > 64 ? 64 : ni; }
// Plane tiles of one slab plane.
__host__ __device__ inline int pml_plane_tiles(int ni, int nj, int cz) {
  if (!pml_tiled(cz)) return (ni * nj + kBlock - 1) / kBlock;
  const int tx = pml_tile_x(ni), ty = kBlock / tx;
  return ((ni + tx - 1) / tx) * ((nj + ty - 1) / ty);
}
// blocks of one CPML slab: plane tiles times chunks of pml_planes(cz) planes
inline int pml_blocks(int ni, int nj, int nk, int cz) {
  const int p = pml_planes(cz);
  return pml_plane_tiles(ni, nj, cz) * ((nk + p - 1) / p);
}

// Pack on store: the kernels that update a subdomain's boundary layers also
// write each updated value into the ghost layers of the neighbour that
// receives it, so no separate CMM exchange pass reads the layers again.
// Side order top, bottom, front, back, right, left (z lo, z hi, y lo, y hi,
// x lo, x hi); a cell at local extended coordinate c along the side's axis
// is sent when c < w (lo sides) or c >= en - w (hi sides), w the width the
// receiver takes for the field (table H1), and lands at the receiver's
// coordinate c + shift (the other two coordinates are equal on both sides).
struct MirrorSide {
  char *base;   // receiver's field 0 at its local extended (0,0,0), bytes; nullptr: no neighbour
  long fpitch;  // elements between the receiver's fields
  long sy, sz;  // receiver's strides
  int shift;
  int w[10];    // width sent per field id VX .. SAU (0: not sent)
};
struct MirrorDesc {
  MirrorSide s[6];
  int en[3];  // the sender's local extended size
};

// One subdomain's physical region (gridphy), in the subdomain's storage.
// Field pointers address local extended cell (0,0,0); strides sy, sz in
// elements; ghost cells toward neighbours live at negative / past-the-end
// local coordinates.  Field elements are of the context's storage format.
struct PhyDesc {
  void *f[NF];
  const float *dtr;  // dt / rho, per cell
  const float *lam;
  const float *mu;
  long sy, sz;
  int ox, oy, oz;    // gridphy origin in local extended coordinates
  int nx, ny, nz;    // gridphy size
  int gx, gy, gz;    // offset of this gridphy inside the global gridphy
  int gnx, gny, gnz; // the global gridphy (VOWR variable order of the backward mode)
  int R;             // VOWR ring width
  int tiles_x, tiles_y;
  float dt;          // time step (FWI accumulators)
  const MirrorDesc *mir;  // pack on store (z-blocked bricks); nullptr: none
  int blk0;
};

// One CPML slab of one subdomain (the reference's top/bottom, front/back,
// right/left PML frames, BASE/fdtd3d_base_6_pml_meat.hpp:29-93).
struct SlabDesc {
  void *f[NF];
  const float *dtr;
  const float *lam;
  const float *mu;
  long sy, sz;
  int i0, j0, k0, ni, nj, nk;        // box in local extended coordinates
  int axes;                          // damped axes: 1 x, 2 y, 4 z
  int xlo, xhi, ylo, yhi, zlo, zhi;  // a cell is in an axis' damping range iff c < lo || c >= hi
  void *psi;                         // [6 * naxes][ncell] memory variables (storage format)
  const float *coef;                 // [4 * naxes][ncell]: a, b, a_half, b_half per axis
  const MirrorDesc *mir;             // pack on store; nullptr: none
  int blk0;
};

// XCD-partitioned block order of the CPML kernels.  Every slab is split
// into 8 parts, one per XCD: by z chunks when the slab is deeper (z) than it
// is long in y, else by plane tiles (kBlock cells of a plane).  Each XCD
// walks its parts chunk by chunk, so a block's z and y stencil taps were
// loaded into the XCD's own L2 by the blocks just before it.  Block b runs
// on XCD b % 8 as that XCD's (b / 8)-th block; a piece covers chunks
// [p0, p0 + n / tcnt) times tiles [t0, t0 + tcnt) of slab `desc`, from
// the XCD's block lstart on.
struct PmlPiece {
  int desc, lstart, p0, t0, tcnt;
};
struct PmlOrder {
  const PmlPiece *pieces = nullptr;  // nullptr: the group-interleaved launch order
  int xoff[9] = {0};                 // pieces of XCD x: [xoff[x], xoff[x + 1])
  int xn[8] = {0};                   // blocks of XCD x
  int grid = 0;                      // 8 x the largest xn
};

// Blocks of the CMM copy of one link: boxes at most kHaloNarrow cells long
// in x (the x-face links, 3-4 cells) take one cell per thread, longer boxes
// one row per wavefront (4 per block).
constexpr int kHaloNarrow = 8;
inline int halo_blocks(int ni, int nj, int nk) {
  const long rows = (long)nj * nk;
  return ni <= kHaloNarrow ? (int)((rows * ni + kBlock - 1) / kBlock)
                           : (int)((rows + kBlock / 64 - 1) / (kBlock / 64));
}

// One CMM halo link of one field: the sender's boundary layers copied
// straight into the receiver's ghost layers (both subdomains live on the
// same device, so the contiguous message block of the reference's MPI
// exchange is not needed).  Same box size on both sides, x fastest.
struct HaloDesc {
  const void *src;
  void *dst;
  long ssy, ssz, dsy, dsz;
  int si, sj, sk;  // box origin in the sender's local coordinates
  int di, dj, dk;  // box origin in the receiver's local coordinates (ghosts)
  int ni, nj, nk;
  int blk0;
};

// One box copied between a field and a contiguous buffer (the VOWR boundary
// store).
struct CopyDesc {
  void *field;
  long sy, sz;
  int i0, j0, k0, ni, nj, nk;
  // element (i, j, k) of the box is buffer element buf_off + (bi + i) +
  // bni * ((bj + j) + bnj * (bk + k)): the box is the sub-box at (bi, bj, bk)
  // of a bni x bnj x .. buffer box (a ring box's part that lies in a
  // neighbour's ghost layers); bi = bj = bk = 0, bni = ni, bnj = nj for a
  // whole box
  int bi, bj, bk, bni, bnj;
  long buf_off;
  int cls;   // field class (0 velocity, 1 stress): the scale of a converting copy
  int fid;   // ring field index (0..9): the per-slice scale of a converting copy
  int blk0;
};

// Imaging over one subdomain's gridphy.
struct ImgDesc {
  const void *b[6];   // backward (source) set: vx vy vz vxp vyp vzp
  const void *r[6];   // reverse (receiver) set
  long sy, sz;
  int ox, oy, oz, nx, ny, nz;
  void *img[9];       // Ip Ipp Ips Isp Iss and 4 Laplacians (float or double)
  int blk0;
};

// Brick of the z-blocked physical-region kernels: a block of tx*ty threads
// owns (px*tx) x (ty*cy) columns and kz consecutive planes (kz per thread);
// px = 2: every thread updates two neighbouring x cells; xl = 1: one cell
// per thread with the x arm taken from neighbouring lanes (tx = 64); xl = 2:
// that, and the y arm from LDS (cy = 1); kz == 0 selects the
// one-plane-per-block kernels.
struct Tiling {
  int tx, ty, cy, kz;
  int px = 1;
  int xl = 0;
};
// instantiated brick shapes (tx, ty, cy, kz) for a storage format
bool tiling_supported(const Tiling &t, int fmt = STORE_F32);
// the z-blocked shapes of a storage format (one-cell, paired, lane-exchange
// and y arm from LDS: xl = 2); returns how many there are and writes up to
// max of them
int tiling_list(int fmt, Tiling *out, int max);
// cy == 0 selects the LDS-tiled 2.5D kernels (tx = 64, ty = 8, kz planes per
// block).
inline int tiling_blocks(const Tiling &t, int nx, int ny, int nz) {
  const int rows = t.cy == 0 ? t.ty : t.ty * t.cy;
  const int cols = t.tx * t.px;
  return ((nx + cols - 1) / cols) * ((ny + rows - 1) / rows) * ((nz + t.kz - 1) / t.kz);
}

// One member of a fused batch of modelling shots: its source cell and
// wavelet, its receivers and seismograms, and the batch step at which its
// shot began (its time index is the batch step minus join).
struct MemberIO {
  void *sxx, *syy, *szz;  // source subdomain's fields; nullptr: the shot is outside gridext
  long src_off;
  const float *wavelet;
  int src_mode;
  float ss;               // stress-class scale (16-bit formats)
  void *const *vptr;      // receivers: per subdomain vx, vy, vz
  const int *rsub;
  const long *roff;
  int nrec;
  float *seis;            // [3][nrec][ntime]
  float inv_sv;
  float sv;               // velocity-class scale (receiver injection of the 16-bit formats)
  int join, ntime;
  int rev;                // 0: time index = batch step - join; 1: join - batch step (a reverse loop)
  int blk0;               // first block of this member's receivers
};

// ---- launchers (kernels.hip) ----
// fmt: storage format; the 16-bit formats fold their scale ratios into
// `sign` (velocity: sv/ss), `dt` (stress: ss/sv) and `rv` (CPML velocity:
// sv/ss); ss, sv, inv_sv: the class scales themselves.
void set_fd_table(const float tab[5][4]);
// block order of the z-blocked bricks: 0 contiguous per XCD, g > 0 groups of g
void set_brick_group(int g);
// The physical-region launchers return false, launching nothing, when no
// kernel exists for the brick shape, storage format and mode.
bool launch_vel_phy(const Tiling &t, int mode, int fmt, const PhyDesc *d, const int *starts, int n, int nblk,
                    float dx, float dy, float dz, float sign, int gnx, int gny, int gnz, int R, hipStream_t st);
bool launch_stress_phy(const Tiling &t, int mode, int fmt, const PhyDesc *d, const int *starts, int n, int nblk,
                       float dx, float dy, float dz, float dt, int gnx, int gny, int gnz, int R, hipStream_t st);
// grp: blocks per XCD group of the group-interleaved block order (0: launch order)
// ord (nullable): the XCD-partitioned block order of the table; without it
// (or with ord->pieces null) the blocks follow the table's own order.
void launch_vel_pml(int cz, int fmt, const SlabDesc *d, const int *starts, int n, int nblk, float dx, float dy,
                    float dz, float rv, int grp, const PmlOrder *ord, hipStream_t st);
void launch_stress_pml(int cz, int with_sau, int fmt, const SlabDesc *d, const int *starts, int n, int nblk,
                       float dx, float dy, float dz, float dt, int grp, const PmlOrder *ord, hipStream_t st);
// dir +1: field -> buffer (VOWR save); -1: buffer -> field (VOWR inject);
// esz: element size in bytes (4 or 2)
void launch_copy(int dir, int esz, const CopyDesc *d, const int *starts, int n, int nblk, void *buf, hipStream_t st);
void launch_halo(int esz, const HaloDesc *d, const int *starts, int n, int nblk, hipStream_t st);
// Converting copy between FP32 fields and a 16-bit buffer (fmt FP16 or
// BF16): saved values are scaled by sc[cls] and rounded; injected ones
// widened and scaled by 1 / sc[cls].
void launch_copy_conv(int dir, int fmt, const CopyDesc *d, const int *starts, int n, int nblk, void *buf, float sc_v,
                      float sc_s, hipStream_t st);
// Per-slice, per-field scaling of a 16-bit VOWR store over FP32 fields:
// launch_ring_max accumulates each field's largest |value| of the boxes into
// maxbits[fid] (float bits; the array starts at zero), launch_ring_exp turns
// them into the slice's scale exponents exps[0..nf) (e = 15 - the binary
// exponent of the maximum, clamped to [-100, 100]; 0 for an all-zero field)
// and clears maxbits again, launch_copy_dyn copies with scale 2^exps[fid]
// (dir +1 save, -1 inject).
void launch_ring_max(const CopyDesc *d, const int *starts, int n, int nblk, unsigned *maxbits, hipStream_t st);
void launch_ring_exp(unsigned *maxbits, int *exps, int nf, hipStream_t st);
void launch_copy_dyn(int dir, int fmt, const CopyDesc *d, const int *starts, int n, int nblk, void *buf,
                     const int *exps, hipStream_t st);
// The block-scaled VOWR store (STORE_BFP16): dir +1 saves the boxes of a
// save table (whole boxes on whole runs) into the slice buf and writes the
// runs' exponents to bexp, dir -1 injects the boxes of any ring table.
void launch_ring_bfp(int dir, const CopyDesc *d, const int *starts, int n, int nblk, void *buf, signed char *bexp,
                     hipStream_t st);
// itp (nullable): read the time index from this device int instead of it
void launch_source(int fmt, void *sxx, void *syy, void *szz, long off, const float *wavelet, int it, float sign,
                   int src_mode, float ss, const int *itp, hipStream_t st);
void launch_advance(int *p, int delta, hipStream_t st);
// Sources and receiver samples of every member of a fused batch at batch
// step gs, or *gsp when gsp is set (a device step counter: launches
// captured once in a graph and replayed every step); members outside
// [join, join + ntime) do nothing; nblk: receiver blocks over all members
// (starts[q] = io[q].blk0).
void launch_source_batch(int fmt, const MemberIO *io, int n, int gs, const int *gsp, hipStream_t st,
                         float sign = 1.f);
// Observed data of every member of a fused RTM batch injected at its
// receivers (io[q].vptr: the reverse set) at batch step gs.
void launch_inject_batch(int fmt, const MemberIO *io, const int *starts, int n, int nblk, int gs, hipStream_t st);
void launch_recv_batch(int fmt, const MemberIO *io, const int *starts, int n, int nblk, int gs, const int *gsp,
                       hipStream_t st);
void launch_set_int(int *p, int v, hipStream_t st);
void launch_recv_sample(int fmt, void *const *vptr, const int *rsub, const long *roff, int nrec, int it, int ntime,
                        float *seis, float inv_sv, const int *itp, hipStream_t st);
void launch_recv_inject(int fmt, void *const *vptr, const int *rsub, const long *roff, int nrec, int it, int ntime,
                        const float *seis, float sv, hipStream_t st);
// v += obs - syn at the receivers (FWI adjoint source; FP32 storage)
void launch_recv_residual(void *const *vptr, const int *rsub, const long *roff, int nrec, int it, int ntime,
                          const float *obs, const float *syn, hipStream_t st);
void launch_image(int fp64, int fmt, const ImgDesc *d, const int *starts, int n, int nblk, float inv_sv,
                  hipStream_t st);
void launch_image_post(int fp64, const ImgDesc *d, const int *starts, int n, int nblk, hipStream_t st);

// Sets *flag to 1 when any of the n elements of a field (storage format
// fmt) is infinite or NaN; leaves it alone otherwise.
void launch_nonfinite(int fmt, const void *field, long n, int *flag, hipStream_t st);

}  // namespace ddx
