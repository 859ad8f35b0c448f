/*
 * ddx.h -- C ABI of the MI355X elastic FDTD / RTM engine (libddx.so).
 *
 * Plain C: opaque context, plain pointers and sizes, int status codes.
 * Every entry point names the reference interface it replaces (paths are
 * relative to the CalvinStudio/domain-decomposition-release root;
 * BASE/ = src/jarvis_cpp/jarvis/.jarvis_2_app/.jarvis_3_seismology/
 * elastic_1_fdtd3d_model_mpi/fdtd3d_0_base_module/, RTM/ = .../
 * elastic_1_rtm3d_model_mpi/rtm3d_1_model_mpi_blend/, etc.).
 *
 * Ownership: the context owns every device buffer (wavefields, CPML state,
 * CMM halo blocks, the VOWR boundary store, seismograms, images).  The caller
 * owns every host buffer passed in or out; nothing is retained after a call.
 * Errors: a nonzero return and a message from ddx_last_error().
 * Threading: one context is driven by one host thread at a time.
 */
#ifndef DDX_H
#define DDX_H

#ifdef __cplusplus
extern "C" {
#endif

#define DDX_OK 0
#define DDX_EINVAL 1   /* bad argument or configuration */
#define DDX_EHIP 2     /* HIP runtime error */
#define DDX_ESTATE 3   /* call out of order (e.g. no model yet) */
#define DDX_ENOMEM 4   /* device memory budget exceeded */
#define DDX_EUNSTABLE 5 /* CFL limit exceeded and policy is refuse */

/* field ids, BASE/fdtd3d_base_2_wavefield_bones.hpp:11-18 (orgWave, rtmWave) */
enum ddx_field {
  DDX_VX = 0, DDX_VY, DDX_VZ, DDX_SXX, DDX_SYY, DDX_SZZ, DDX_SXY, DDX_SXZ, DDX_SYZ,
  DDX_SAU, DDX_VXP, DDX_VYP, DDX_VZP,
  /* FWI accumulators on gridphy (fwiWave, BASE/fdtd3d_base_2_wavefield_bones.hpp:19-22) */
  DDX_UX, DDX_UY, DDX_UZ, DDX_VXRT, DDX_VYRT, DDX_VZRT, DDX_NFIELD
};

/* propagation modes, BASE/fdtd3d_base_1_interface.hpp:19-26 (SimulateType) */
enum ddx_mode { DDX_PURE_FORWARD = 0, DDX_RTM_FORWARD = 1, DDX_RTM_BACKWARD = 2, DDX_RTM_REVERSE = 3, DDX_FWI = 4 };

/* image ids, IMG/rtm3d_image_bones.hpp:18-28 (MigImage members) */
enum ddx_image {
  DDX_IP = 0, DDX_IPP, DDX_IPS, DDX_ISP, DDX_ISS,
  DDX_IPP_LAPLACE, DDX_IPS_LAPLACE, DDX_ISP_LAPLACE, DDX_ISS_LAPLACE, DDX_NIMAGE
};

/* Run description.  Replaces ArgList (BASE/fdtd3d_base_1_interface.hpp:27-113)
 * plus GeoConst (:10-16): phy_fdorder_half = 4 and pml_fdorder_half = 2 are
 * fixed by the kernels; rec_width is a run-time value (1..4). */
typedef struct ddx_desc {
  int nx, ny, nz;            /* global physical grid (gridphy) */
  float dx, dy, dz;          /* spacing in metres */
  int pml_num;               /* margin cells per side, including 2 rigid layers */
  int x_divide_num, y_divide_num, z_divide_num; /* subdomains, all on one GPU */
  float dt;                  /* time step in seconds */
  float fm;                  /* dominant frequency (CPML alpha_max = pi*fm) */
  int ntime;                 /* number of time steps */
  int rec_width;             /* VOWR ring width R, 1..4 */
  int image_fp64;            /* 1: accumulate RTM images in double */
  int source_mode;           /* 0: explosive (sxx,syy,szz), 1: szz only */
  int device;                /* HIP device ordinal */
  int cfl_policy;            /* 0: refuse an unstable model, 1: warn only */
  int storage;               /* wavefield storage format (enum ddx_storage); 0 = FP32 as the reference */
  int ring_storage;          /* VOWR boundary store format over FP32 wavefields: 0 = as `storage`,
                                DDX_STORE_FP16 or DDX_STORE_BF16 (values rounded when saved),
                                or DDX_STORE_BFP16 (16-bit codes with per-run exponents) */
} ddx_desc;

/* Wavefield storage formats.  Every wavefield, CPML memory variable and the
 * VOWR boundary store hold FP32 (the reference's type), or 16-bit codes of the
 * value times a power-of-two scale of its field class; stencils, sources,
 * receivers and images compute in FP32 (images optionally FP64).  No
 * counterpart in the reference.  The 16-bit formats run the z-blocked brick
 * kernels and not the FWI mode. */
enum ddx_storage { DDX_STORE_FP32 = 0, DDX_STORE_FP16 = 1, DDX_STORE_BF16 = 2, DDX_STORE_BFP16 = 3 };
/* DDX_STORE_BFP16, for ring_storage only: each run of 16 consecutive entries
 * of a VOWR time slice shares one exponent s (one byte; the run's largest
 * |value| is below 2^s), and each entry holds a sign, a 3-bit exponent and a
 * 12-bit mantissa of |value| 2^-s.  Values within 2^-7 of the run's largest
 * keep a relative rounding error of at most 2^-13 (FP16: 2^-11); smaller
 * ones an absolute error of at most 2^(s-20).  2 bytes per value plus 1/16
 * byte. */
/* field classes of the storage scales: velocity class (vx..vz, vxp..vzp and
 * the memory variables of the stress update), stress class (sxx..syz, sau and
 * the memory variables of the velocity update) */
enum ddx_field_class { DDX_CLASS_VELOCITY = 0, DDX_CLASS_STRESS = 1 };

typedef struct ddx_ctx ddx_ctx;

/* ---- context -------------------------------------------------------------
 * Replaces the module wiring of rtm3d_blend_model_mpi_module::link
 * (RTM/rtm3d_blend_elastic_model_mpi_meat.hpp:6-16) and the per-rank
 * decomposition elasticGeoModel<local>::set_gridphy/set_gridext
 * (BASE/fdtd3d_base_3_geomodel_meat.hpp:579-704): all subdomains of the
 * x*y*z decomposition are created on one device. */
int ddx_create(const ddx_desc *desc, ddx_ctx **out);
void ddx_destroy(ddx_ctx *ctx);
/* Last error message of ctx, or of the last failed ddx_create when ctx is NULL. */
const char *ddx_last_error(const ddx_ctx *ctx);

/* ---- inputs --------------------------------------------------------------
 * Model on the global physical grid, float32, x fastest (nx*ny*nz each).
 * Replaces elasticGeoModel<global>::initialize + extend_model + lambda/mu
 * (BASE/fdtd3d_base_3_geomodel_meat.hpp:100-364, :405-460),
 * elasticGeoModel<local>::set_model_para (:705-718) and the CPML damping
 * coefficients mpicuCPML::calculate_damp_coeff
 * (CPML/fdtd3d_cpml_mpi_1_host_meat.hpp:147-620). */
int ddx_set_model(ddx_ctx *ctx, const float *vp, const float *vs, const float *rho);
/* Optional migration model (same layout): when set, the RTM passes
 * (forward with boundary save, backward reconstruction, receiver
 * back-propagation) run on it while ddx_model_shot keeps the model of
 * ddx_set_model, as the smoothed copies lambda_smo / mu_smo / rho_smo of
 * elasticGeoModel<global>::initialize intend
 * (BASE/fdtd3d_base_3_geomodel_meat.hpp:466-474; computed but unused
 * there).  Three NULL pointers remove it. */
int ddx_set_migration_model(ddx_ctx *ctx, const float *vp, const float *vs, const float *rho);
/* Source wavelet, ntime samples (SeisRecord::wave_let,
 * BASE/fdtd3d_base_4_seisrecord_meat.hpp:273-297). */
int ddx_set_wavelet(ddx_ctx *ctx, const float *wavelet);
/* Shots and receivers in metres (x,y,z triples), as read by
 * SubGeometry::initialize (UTIL/.jarvis_2_geometry/jarvis_device_geometry_meat.hpp:59-128).
 * Receivers outside the physical grid are dropped; kept receivers are
 * ordered by subdomain rank, then by input order, as the per-rank files. */
int ddx_set_geometry(ddx_ctx *ctx, const float *shots_xyz, int nshot, const float *recv_xyz, int nrecv);
/* Number of kept receivers, and for each kept row its subdomain rank and
 * input index (either pointer may be NULL). */
int ddx_receivers(const ddx_ctx *ctx, int *nkept, int *rank_of_row, int *input_of_row);

/* ---- drivers ---------------------------------------------------------------
 * Pure forward modelling of one shot; the seismograms stay on the device as
 * that shot's observed data.  Replaces
 * fdtd3d_blend_elastic_model_mpi_module<pure_forward>::mpi_sub_forward
 * (BLEND/fdtd3d_blend_elastic_model_mpi_meat.hpp:44-88). */
int ddx_model_shot(ddx_ctx *ctx, int ishot);
/* Seismograms of a shot to host, [3][nkept][ntime] (vx, vy, vz; time fastest),
 * SeisRecord::copy_seis_into_host (BASE/..._4_seisrecord_meat.hpp:299-312). */
int ddx_get_seismograms(ddx_ctx *ctx, int ishot, float *out);
/* Observed data of a shot from host, same layout (copy_seis_from_host, :313-324). */
int ddx_set_observed(ddx_ctx *ctx, int ishot, const float *in);
/* One RTM shot: forward with VOWR boundary save, then the reverse loop of
 * backward reconstruction, receiver back-propagation and imaging.  Images
 * accumulate over shots.  Replaces mpi_sub_forward_for_restruct
 * (BLEND/...meat.hpp:125-155) and the reverse loop of mpi_sub_rtm3d
 * (RTM/rtm3d_blend_elastic_model_mpi_meat.hpp:50-122). */
int ddx_rtm_shot(ddx_ctx *ctx, int ishot);
/* Shots in pieces (no reference counterpart): stages s0 .. s0+n-1 of shot
 * ishot, for a modelling shot its ntime time steps, for an RTM shot 2*ntime
 * stages (the forward steps with boundary save, then the reverse iterations
 * ntime-1 .. 0).  A piece starting at stage 0 starts the shot (as
 * ddx_model_shot / ddx_rtm_shot do); every other piece must continue where
 * the previous one stopped (DDX_ESTATE otherwise).  All pieces together
 * equal the whole-shot call; between pieces a caller (a serving loop) may
 * check the wavefields or stop the shot. */
int ddx_model_steps(ddx_ctx *ctx, int ishot, int s0, int n);
int ddx_rtm_steps(ddx_ctx *ctx, int ishot, int s0, int n);
/* *finite = 0 when any velocity of either wavefield set holds an infinity
 * or a NaN (a run that diverged), 1 otherwise.  No reference counterpart
 * (the reference's stability check is disabled,
 * PHY/fdtd3d_phy_mpi_1_host_meat.hpp:150-184). */
int ddx_check_finite(ddx_ctx *ctx, int *finite);
/* Laplacian then source compensation (RTM/...meat.hpp:123-124;
 * IMG/rtm3d_iamge_kernel.hpp:118-135, :184-200). */
int ddx_finalize_images(ddx_ctx *ctx);
int ddx_reset_images(ddx_ctx *ctx);
/* One image of one subdomain (rank), on that subdomain's gridphy, x fastest,
 * as written per rank by RTM/...meat.hpp:126-134. */
int ddx_get_image(ddx_ctx *ctx, int which, int rank, float *out);

/* Snapshots of one field of the source-side set during ddx_rtm_shot, kept
 * on the device: every `every` steps of the forward pass and the matching
 * steps of the backward reconstruction (no reference counterpart; it checks
 * the VOWR reconstruction against the stored full source wavefield).  Index
 * q holds the forward state after step ntime-1-q*every and the backward
 * state after reverse iteration ntime-q*every (q >= 1 for the backward
 * pass), on the global gridphy, x fastest.  every = 0 frees them. */
int ddx_set_snapshots(ddx_ctx *ctx, int field, int every);
int ddx_get_snapshot(ddx_ctx *ctx, int pass, int index, float *out);

/* ---- step-level hooks (parity checks, custom drivers) ----------------------
 * set 0 is the forward/backward wavefield set, set 1 the reverse set. */
int ddx_reset_wave(ddx_ctx *ctx, int set);
/* Velocity (resp. stress) half step of every subdomain, then the CMM halo
 * exchange of that batch (mpicuPhy::cuda_update_and_mpi_exchange_edge_vel,
 * PHY/fdtd3d_phy_mpi_4_device_meat.hpp:142-162, with the CPML halves
 * CPML/fdtd3d_cpml_mpi_3_device_meat.hpp:7-579). */
int ddx_step_vel(ddx_ctx *ctx, int set, int mode);
int ddx_step_stress(ddx_ctx *ctx, int set, int mode);
/* The same half steps without the exchange (cuda_update_inside_vel/_stress
 * plus the edge and CPML updates, PHY/fdtd3d_phy_mpi_4_device_meat.hpp:
 * 163-268), and the CMM exchange of one batch on its own
 * (mpicu_halo_comm_Operator::extract_into_halo + the MPI exchange,
 * OP/jarvis_2_mpicu_halo_comm_operator_meat.hpp:299-326): batch 0 the
 * velocities, 1 the stresses, 2 the stresses and sau.  Every subdomain is
 * on the one device, so the exchange copies each sender's boundary layers
 * into the receiver's ghost layers. */
int ddx_update_vel(ddx_ctx *ctx, int set, int mode);
int ddx_update_stress(ddx_ctx *ctx, int set, int mode);
int ddx_halo_exchange(ddx_ctx *ctx, int set, int batch);
/* How the CMM exchange moves the halos (the reference exchanges them by
 * mpicu_halo_comm_Operator::extract_into_halo / inject_from_halo,
 * OP/jarvis_2_mpicu_halo_comm_operator_meat.hpp:299-326).  mode 0 (the
 * default): one copy per link from the sender's boundary layers straight
 * into the receiver's ghost layers.  mode 1: the reference's contiguous
 * message blocks -- every sender side packs the boundary layers of the
 * batch's fields into one contiguous block, each field at its offset in
 * push order (the blocks of the multiple-halo operator, :408-461), and the
 * receiver unpacks the block into its ghost layers; the block is what a
 * link between two GPUs would send.  mode 2: pack on store -- the kernels
 * that update a subdomain's boundary layers (z-blocked bricks, CPML
 * kernels, VOWR injection) also write each value into the receiver's ghost
 * layers, so no pass reads the layers again; after a half step run by the
 * one-plane or LDS-tiled kernels the mode-0 copy follows instead.  Mode 2
 * needs the context created with DDX_MIRROR=1 in the environment (it is
 * then the context's default), and is slower on this GPU than mode 0.
 * Results do not depend on the mode.  In mode 2, ddx_update_vel /
 * ddx_update_stress already leave the ghosts current, and ddx_halo_exchange
 * copies them again (same values). */
int ddx_set_exchange(ddx_ctx *ctx, int mode);
int ddx_get_exchange(const ddx_ctx *ctx, int *mode);
/* VOWR boundary store of set 0: save time slice it (0..ntime-2) of every
 * ring field (elasticWaveRecon::extract_boundary_wave +
 * copy_into_host_storage, BASE/fdtd3d_base_5_wave_recon_meat.hpp:60-88; the
 * store stays in HBM), and inject slice it of the velocities (which = 1) or
 * stresses (which = 2) (inject_boundary_vel/_stress, :89-133). */
int ddx_vowr_save(ddx_ctx *ctx, int it);
int ddx_vowr_inject(ddx_ctx *ctx, int which, int it);
/* Observed data of shot ishot at sample it added to the velocities of `set`
 * at the receivers (sub_rtm_reverse_add_vel_source,
 * BASE/fdtd3d_base_4_seisrecord_meat.hpp:79-107). */
int ddx_recv_inject(ddx_ctx *ctx, int set, int ishot, int it);
/* One zero-lag cross-correlation of the backward (set 0) and reverse (set 1)
 * wavefields into the images (cal_correlation_from_wave,
 * IMG/rtm3d_iamge_kernel.hpp:6-51). */
int ddx_image_step(ddx_ctx *ctx);
/* FWI hooks.  In mode DDX_FWI the velocity half step also accumulates
 * u += dt v and stores vrt = dt x the velocity increment on gridphy
 * (PHY/fdtd3d_phy_mpi_2_vel_kernel.hpp:70-78); the stress half step is the
 * pure-forward one.  ddx_record_synthetic samples the receivers of `set`
 * at step `it` into the shot's synthetic data; ddx_inject_residual adds
 * observed - synthetic at the receivers (the fwi branch of
 * sub_add_reverse_vel_source, BASE/fdtd3d_base_4_seisrecord_meat.hpp:109-147). */
int ddx_record_synthetic(ddx_ctx *ctx, int set, int ishot, int it);
int ddx_get_synthetic(ddx_ctx *ctx, int ishot, float *out);
int ddx_inject_residual(ddx_ctx *ctx, int set, int ishot, int it);
/* Source add (sign +1) or subtract (-1) of sample `it` of shot `ishot`
 * (PHY/fdtd3d_phy_mpi_3_stress_kernel.hpp:1308-1364). */
int ddx_source(ddx_ctx *ctx, int set, int ishot, int it, int sign);
/* Whole field on the global extended grid (nx+2*pml etc., x fastest),
 * assembled from the subdomains; cells no subdomain owns read as 0.  With a
 * 16-bit storage format the values are the stored ones, unscaled; set_field
 * rounds to the format as every store does. */
int ddx_get_field(ddx_ctx *ctx, int set, int field, float *out);
int ddx_set_field(ddx_ctx *ctx, int set, int field, const float *in);
int ddx_synchronize(ddx_ctx *ctx);

/* Brick shape of the z-blocked physical-region kernels (no reference
 * counterpart; the reference launches one thread per cell in 1D blocks of
 * 128, MATH/jarvis_2_device_math_bones.hpp:14): a block of tx*ty threads
 * owns tx x (ty*cy) columns and kz consecutive planes; kz = 0 selects one
 * plane per block.  Supported (tx,ty,cy,kz): (64,4,1,2) (64,4,1,4) (64,4,1,8)
 * (128,2,1,4) (64,8,1,4) (32,8,1,4) (64,4,2,4) (128,4,1,2) (64,2,1,8), and
 * (64,4,1,0); (64,8,0,kz) with any kz > 0 selects the LDS-tiled kernels
 * (64x8 columns staged in LDS per plane, kz planes per block).  The 16-bit
 * storage formats take (64,4,1,2) (64,4,1,4) (64,4,1,8) (128,2,1,4)
 * (64,2,1,8).  Results do not depend on the brick shape. */
int ddx_set_tiling(ddx_ctx *ctx, int tx, int ty, int cy, int kz);
/* Brick shape for one kernel (0 velocity, 1 stress) in one mode only
 * (same shapes as ddx_set_tiling; no reference counterpart). */
int ddx_set_tiling_mode(ddx_ctx *ctx, int kernel, int mode, int tx, int ty, int cy, int kz);
/* Brick shape including its lane form px: 1 one cell per thread; 2 the
 * paired bricks (two neighbouring x cells per thread, one 2-element access
 * per field), shapes (tx,ty,cy,kz) (64,4,1,1) (64,4,1,2) (64,4,1,4)
 * (64,2,1,4) in every storage format; 3 one cell per thread with the x
 * stencil arm taken from neighbouring lanes of the wavefront (DPP shifts
 * instead of loads), shapes (64,4,1,2) (64,4,1,4) (64,4,1,8) (64,8,1,4)
 * (64,2,1,8), FP32 storage; 4 form 3 with the y arm read from an LDS tile
 * of the block's rows, shapes (64,ty,1,kz) for ty in {4, 8, 16} and kz in
 * {2, 4}, FP32 storage.  Forms 2 to 4 do not run the FWI mode.
 * ddx_get_brick returns the shape in use as tx, ty, cy, kz, px.  No
 * reference counterpart. */
int ddx_set_brick(ddx_ctx *ctx, int kernel, int mode, int tx, int ty, int cy, int kz, int px);
int ddx_get_brick(const ddx_ctx *ctx, int kernel, int mode, int *out5);
/* Brick autotuner (no reference counterpart): for each physical-region
 * kernel and propagation mode, times `reps` launches of every z-blocked
 * brick shape of the storage format on this context and keeps the fastest;
 * cached per process by storage format, decomposition and subdomain size
 * (rounded up to 32 cells per axis).  Needs the model; resets both
 * wavefield sets.  *timed (nullable): shapes timed (0 when all came from the
 * cache). */
int ddx_autotune(ddx_ctx *ctx, int reps, int *timed);
/* Modelling shots of several contexts in one launch sequence: shot
 * ishots[q] of ctxs[q] for q < n.  The contexts must be on one device with
 * equal dt, grid spacing, ntime and storage format (and equal scales), and
 * be ready for ddx_model_shot.  Each time step runs one launch of the
 * physical-region velocity and stress bricks and one of each CPML kernel
 * over the cells of every context (brick shapes of ctxs[0]), then each
 * context's halo exchange, source and receivers, on ctxs[0]'s stream.
 * Every context's seismograms equal those of its own ddx_model_shot.  The
 * call owns all the contexts until it returns; errors are reported on
 * ctxs[0].  No reference counterpart (the reference runs one shot at a
 * time). */
int ddx_model_batch(ddx_ctx *const *ctxs, const int *ishots, int n);

/* ddx_model_batch that allocates no device memory.  No counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * fused descriptor tables, which ddx_model_batch allocates per call, go
 * into a fixed device array of 4 MiB per GPU instead (DDX_EINVAL when a
 * batch's tables need more).  Every member must be ready without
 * allocating: its shot's seismogram buffer exists (ddx_model_shot or
 * ddx_set_observed ran on it once) and its tables of ctxs[0]'s brick
 * shapes are built (one shot ran with them); DDX_ESTATE otherwise.  Runs
 * the time steps eagerly (no graph replay).  Seismograms equal those of
 * ddx_model_batch.  The array is shared by every call on one GPU and keeps
 * the last call's tables: calls that overlap on one GPU (from several host
 * threads) race and must be serialised by the caller. */
int ddx_model_batch_noalloc(ddx_ctx *const *ctxs, const int *ishots, int n);

/* RTM shots of several contexts in one launch sequence: shot ishots[q] of
 * ctxs[q] (observed data set, as for ddx_rtm_shot) for q < n.  The contexts
 * share what ddx_model_batch requires plus ntime and the image precision;
 * grids, CPML widths, ring widths, models (and migration models),
 * wavelets, receivers and decompositions may differ.  Each time step of the
 * forward passes and each reverse iteration runs one launch per
 * physical-region, CPML and imaging kernel over every member's cells (brick
 * shapes of ctxs[0]); each member's images then equal those of its own
 * ddx_rtm_shot.  Replaces the reference's sequential shot loop,
 * RTM/rtm3d_blend_elastic_model_mpi_meat.hpp:50-122, for shots that fit the
 * GPU together. */
int ddx_rtm_batch(ddx_ctx *const *ctxs, const int *ishots, int n);

/* Continuous batch of modelling shots (no reference counterpart): members
 * join with ddx_batch_add at any step boundary and leave when their shot
 * has run its ntime steps.  Each ddx_batch_run step launches the
 * physical-region bricks and the CPML kernels once over every member's
 * cells (brick shapes of the first member ever added), then each member's
 * halo exchange, source and receivers at the member's own time index, on
 * the batch's own stream.  Members share what ddx_model_batch requires
 * except ntime.  A member's seismograms equal those of its ddx_model_shot.
 * The batch owns a member context from ddx_batch_add until ddx_batch_run
 * returns it: call nothing else on it, and do not destroy it, meanwhile. */
typedef struct ddx_batch ddx_batch;
int ddx_batch_create(ddx_batch **out);
/* Add shot ishot of c: zeroes the context's wavefields; the shot starts at
 * time index 0 in the next ddx_batch_run. */
int ddx_batch_add(ddx_batch *b, ddx_ctx *c, int ishot);
/* Add RTM shot ishot of c (its observed data set, as for ddx_rtm_shot): both
 * wavefield sets are zeroed; from the next batch step it runs its ntime
 * forward steps with VOWR boundary save, then its ntime reverse iterations
 * (backward reconstruction, receiver back-propagation, imaging), and
 * ddx_batch_run returns it when its images hold the shot.  Each step
 * launches the physical-region, CPML and imaging kernels once per group
 * (modelling members, RTM members in their forward pass, RTM members in
 * their reverse loop) over the group's cells.  RTM members share the image
 * precision of the first one.  A member's images equal those of its own
 * ddx_rtm_shot.  Replaces the reference's shot loop
 * (RTM/rtm3d_blend_elastic_model_mpi_meat.hpp:50-122) for shots that arrive
 * while others run. */
int ddx_batch_add_rtm(ddx_batch *b, ddx_ctx *c, int ishot);
/* Advance every member up to max_steps steps, stopping at the first step
 * after which one or more members have finished; writes up to cap finished
 * contexts (their seismograms complete) to done and their count to *ndone.
 * Finished contexts beyond cap are returned by the next call, which then
 * runs no step. */
int ddx_batch_run(ddx_batch *b, int max_steps, ddx_ctx **done, int cap, int *ndone);
int ddx_batch_size(const ddx_batch *b, int *members);
/* Take member c out before its shot has ended (a cancelled or diverged
 * request): it keeps its partial state and is the caller's again. */
int ddx_batch_remove(ddx_batch *b, ddx_ctx *c);
const char *ddx_batch_last_error(const ddx_batch *b);
/* Members still in the batch keep partial wavefields and seismograms. */
int ddx_batch_destroy(ddx_batch *b);
/* Graph replay (no reference counterpart): with on = 1, ddx_model_shot
 * captures one time step (velocity, source, stress, receivers, with the CMM
 * exchanges) in a HIP graph and replays it every step; the time index comes
 * from a device counter that the step advances.  Cuts the host launch cost
 * of small shots; results are identical.  Off by default, and skipped while
 * ddx_profile is timing a kernel class. */
int ddx_set_graphs(ddx_ctx *ctx, int on);

/* ---- measurement ------------------------------------------------------------
 * Time every launch of one kernel class with HIP events on the context
 * stream.  kernel: 0 velocity (physical region), 1 stress (physical region),
 * 2 CPML velocity, 3 CPML stress, 4 CMM halo exchange, 5 imaging; -1 every
 * class at once; any other value stops timing.
 * ddx_kernel_time returns the summed milliseconds and the launch count since
 * the last ddx_profile call, and the algorithmic bytes moved by those
 * launches (every class's, summed, after ddx_profile(ctx, -1)). */
int ddx_profile(ddx_ctx *ctx, int kernel);
int ddx_kernel_time(ddx_ctx *ctx, double *ms, long long *launches, double *bytes);
/* With ddx_profile(ctx, -1) every class is timed at once; this returns one
 * class's figures (valid in either mode).  No reference counterpart. */
int ddx_kernel_class_time(ddx_ctx *ctx, int kernel, double *ms, long long *launches, double *bytes);
/* Cell updates performed since creation (Σ cells × steps over passes). */
int ddx_cell_updates(const ddx_ctx *ctx, double *cells);
/* Device bytes held by the context. */
int ddx_device_bytes(const ddx_ctx *ctx, double *bytes);

/* ---- host-only helpers (no GPU needed) ---------------------------------------
 * Subdomain geometry of `rank`: out[0..2] phy n, [3..5] global phy offset,
 * [6..8] local ext n, [9..14] margins (top,bottom,front,back,right,left),
 * [15..20] neighbour ranks in the same side order (-1: none).
 * BASE/fdtd3d_base_3_geomodel_meat.hpp:579-704 and
 * OP/jarvis_2_mpicu_halo_comm_operator_meat.hpp:26-101. */
int ddx_subdomain(const ddx_desc *desc, int rank, int *out21);
/* Box smoothing of an nx*ny*nz grid (x fastest) along x, y, z with
 * half-widths rx, ry, rz (windows clipped to the grid, mean in double), the
 * smoothing of BASE/fdtd3d_base_3_geomodel_meat.hpp:7-80 with its y and z
 * passes working.  in and out may alias. */
int ddx_smooth_model(const float *in, float *out, int nx, int ny, int nz, int rx, int ry, int rz);
/* Device bytes a context for `desc` will hold with nshot shots of nrec
 * kept receivers (rtm: include the VOWR store; migration: a second model),
 * for admission control (no reference counterpart; the reference only
 * aborts below 10 GB free, MATH/jarvis_2_device_math_meat.hpp:60-64). */
int ddx_estimate_bytes(const ddx_desc *desc, int nrec, int nshot, int rtm, int migration, double *bytes);

/* Storage scale of a field class: fields of the class hold x * 2^log2_scale
 * in the 16-bit formats (|log2_scale| <= 60); with FP32 wavefields and a
 * 16-bit ring_storage, the scale of the VOWR store's values.  Defaults: BF16 0; FP16 chosen
 * from the wavelet peak and the model when they are set (the source cell's
 * stress at about 2^10, velocities at the stress over the mean rho*vp), kept
 * once set here.  Set between shots: it applies to values stored afterwards.
 * No counterpart in the reference (FP32 storage needs no scale). */
int ddx_set_storage_scale(ddx_ctx *c, int field_class, int log2_scale);
int ddx_get_storage_scale(const ddx_ctx *c, int field_class, int *log2_scale);
/* Scaling of a 16-bit VOWR store over FP32 wavefields (no counterpart in
 * the reference, whose store is FP32): mode 0 (default) the class scales
 * above; mode 1 one power-of-two scale per ring field and time slice, set
 * when the slice is saved so that the field's largest |value| in it lands
 * in [2^14, 2^15) (the exponents are kept beside the store).  Small
 * boundary values then keep the format's full relative precision. */
int ddx_set_ring_scaling(ddx_ctx *c, int mode);
/* Staggered FD coefficients of half order `half` (1..4), in float
 * (UTIL/.jarvis_2_geometry/jarvis_device_geometry_bones.hpp:29-49). */
int ddx_fd_coef(int half, float *out);
/* ntime = ceil(T/dt) and the Ricker wavelet of SeisRecord::initialize
 * (BASE/fdtd3d_base_4_seisrecord_meat.hpp:273-285). */
int ddx_ntime(float T, float dt);
int ddx_ricker(float fm, float dt, int ntime, float *out);
/* Largest stable dt for the 8th-order scheme on this model
 * (PHY/fdtd3d_phy_mpi_1_host_meat.hpp:142-148, dt_cal). */
float ddx_stable_dt(float vmax, float dx, float dy, float dz);

#ifdef __cplusplus
}
#endif
#endif /* DDX_H */
