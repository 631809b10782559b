/*
 * surfatt_eik.h -- C ABI of the MI355X (gfx950) eikonal forward + adjoint path.
 *
 * Plain C, POD arguments only: no HIP, torch or Eigen types.  Library:
 * surfattpp_amd/libsurfatt_eik.so.
 *
 * Field layout everywhere: a 2-D field on an nx-by-ny lon/lat grid is a
 * column-major (ix fastest) array of nx*ny doubles, element (ix, iy) at
 * ix + iy*nx -- the storage of the reference's Eigen::MatrixX<real_t>
 * (src/eikonal_solver.cpp, Eigen default) and of its Fortran twin.
 * Per-period planes in the batched API are period-major: plane iper starts
 * at iper*nx*ny (the reference keeps [ix][iy][iper], include/surf_grid.h:63-65;
 * the C++ shim transposes once per period).
 *
 * Status codes: SAT_OK (0) or a positive errno-style value.  Warnings of the
 * reference (non-convergence, src/eikonal_solver.cpp:304-313; source on the
 * boundary, :123-125) are reported as SAT_FLAG_* bits in sat_info.flags and
 * never abort.
 */
#ifndef SURFATT_EIK_H
#define SURFATT_EIK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SAT_OK 0
#define SAT_EINVAL 22  /* shape < 3x3, non-uniform axis, null pointer, bad descriptor */
#define SAT_ENOMEM 12  /* device allocation failed */
#define SAT_EDEVICE 5  /* HIP runtime / kernel launch error, or no gfx950 device */

/* Arithmetic of the sweeps.  SAT_PREC_F64 reproduces the reference's FP64
 * operation order (kernels within 1e-12 of the FP64 oracle).  SAT_PREC_F32
 * sweeps the factored perturbation in FP32 with FP64 source initialisation,
 * convergence norm, misfit and kernel accumulation.  SAT_PREC_F32P (batched
 * runs only) is SAT_PREC_F32 where the forward node update of isotropic grids
 * finds its candidate in FP32 and then takes that candidate's root in FP64
 * from FP64 inputs (one Newton step): travel times ~30x closer to FP64 and
 * kernels ~4x, for a slower forward sweep.  Grids with a topography metric,
 * per-solve runs (SAT_RUN_PER_SOLVE) and sat_eik_forward/backward treat it as
 * SAT_PREC_F32; its batched sweeps use one wave per pair. */
#define SAT_PREC_F64 0
#define SAT_PREC_F32 1
#define SAT_PREC_F32P 2

#define SAT_FLAG_NOT_CONVERGED 1   /* forward hit MAXITER (eikonal_solver.cpp:304) */
#define SAT_FLAG_SRC_ON_BOUNDARY 2 /* a fixed near-source node is on the edge (:123) */
#define SAT_FLAG_ADJ_MAXITER 4     /* adjoint hit maxiter (silent in the reference, :471) */
#define SAT_FLAG_REC_OUTSIDE 8     /* a receiver lies outside the grid: syn = NaN (utils.h:379) */

typedef struct sat_info {
  int iters;     /* sweeping iterations performed (4 sweeps each) */
  int flags;     /* SAT_FLAG_* */
  double l1;     /* last convergence L1 (mean |dtau|*T0 or |dTa|) */
  double linf;   /* last convergence Linf */
} sat_info;

/* ------------------------------------------------------------------------
 * Per-solve calls (host pointers, synchronous, one solve per launch).
 *
 * Grid sizes, here and in the batches: nx >= 3 and ny >= 3 as in the
 * reference; nx <= 4096 (grids wider than 512 columns sweep with 2, 4 or 8
 * waves per pair; wider than 4096: SAT_EINVAL).
 * ------------------------------------------------------------------------ */

/* Replaces eikonal::FSM_UW_PS_lonlat_2d (include/eikonal_solver.h:32-40,
 * src/eikonal_solver.cpp:22-318).  lon_deg[nx], lat_deg[ny] uniformly spaced
 * degrees; spha/sphb/sphc/fun nx*ny; T_out nx*ny travel time (s). */
int sat_eik_forward(int precision, int nx, int ny, const double* lon_deg, const double* lat_deg,
                    const double* spha, const double* sphb, const double* sphc, const double* fun,
                    double src_lon_deg, double src_lat_deg, double* T_out, sat_info* info);

/* sat_eik_forward for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments and results as
 * sat_eik_forward, but lon_deg, lat_deg, spha, sphb, sphc, fun and T_out are
 * DEVICE pointers (the caller allocates them); info stays a host pointer.
 * Calls no hipMalloc/hipFree: the planes of the solve live in one static
 * device array of 24 MiB per GPU, held until the process ends (enough for
 * about 130,000 nodes in FP64, e.g. 360x360; a larger grid returns SAT_EINVAL
 * with a message on stderr -- use sat_eik_forward).  Synchronous on the null
 * stream.  Every call on one GPU shares that array: calls that overlap on one
 * GPU (several host threads) race and must be serialised by the caller. */
int sat_eik_forward_noalloc(int precision, int nx, int ny, const double* lon_deg, const double* lat_deg,
                            const double* spha, const double* sphb, const double* sphc, const double* fun,
                            double src_lon_deg, double src_lat_deg, double* T_out, sat_info* info);

/* Replaces eikonal::FSM_O1_JSE_lonlat_2d (include/eikonal_solver.h:74-83,
 * src/eikonal_solver.cpp:323-512): the adjoint-state field for receiver
 * adjoint sources adj[nr] at (rec_lon_deg, rec_lat_deg), given the forward
 * travel time T (nx*ny). */
int sat_eik_backward(int precision, int nx, int ny, const double* lon_deg, const double* lat_deg,
                     const double* spha, const double* sphb, const double* sphc, const double* T,
                     int nr, const double* rec_lon_deg, const double* rec_lat_deg, const double* adj,
                     double* Ta_out, sat_info* info);

/* Replaces eikonal::mask_uniform_grid (include/eikonal_solver.h:85-89,
 * src/eikonal_solver.cpp:514-536), in place on a host field.  Coordinates in
 * the caller's units (degrees at src/preproc.cpp:125). */
int sat_eik_mask(int nx, int ny, const double* xx, const double* yy, double* Ta, double x0, double y0);

/* ------------------------------------------------------------------------
 * Batched path: every (source, period) pair of one inversion iteration in one
 * persistent launch.  Replaces the per-rank serial loop of
 * preproc::forward_for_event (include/preproc.h:14, src/preproc.cpp:71-148)
 * and its MPI round-robin over events (include/parallel.h:39-41).
 * ------------------------------------------------------------------------ */

typedef struct sat_grid_desc {
  int nx, ny;            /* nodes (>= 3 each) */
  int nperiod;           /* planes below */
  const double* lon_deg; /* [nx] uniform, ascending */
  const double* lat_deg; /* [ny] uniform, ascending */
  const double* svel;    /* [nperiod][ny][nx] phase velocity (km/s) */
  const double* m11;     /* [nperiod][ny][nx] metric (surf_grid.cpp:184-190) */
  const double* m12;     /* [nperiod][ny][nx]; the solver is called with sphc = -m12 */
  const double* m22;     /* [nperiod][ny][nx] */
  int waves;             /* waves (64 lanes each) per pair in the batched sweeps of FP32
                            batches: 1, 2 or 4; 0: chosen by the library (more waves for the
                            largest grids while the batch stays within three waves per SIMD).
                            FP64 batches: 0 or 1.  Results do not depend on it. */
} sat_grid_desc;

typedef struct sat_event_desc {
  int grid;           /* index into grids */
  int iper;           /* period plane */
  double src_lon_deg; /* evlo */
  double src_lat_deg; /* evla */
  int rec_off;        /* first receiver in the flattened receiver arrays */
  int nrec;           /* receivers of this event (may be 0) */
} sat_event_desc;

typedef struct sat_batch_desc {
  int ngrid;
  const sat_grid_desc* grids;
  int nevent;
  const sat_event_desc* events; /* accumulation and chi order = this order */
  int nrec_total;
  const double* rec_lon_deg;    /* [nrec_total] */
  const double* rec_lat_deg;
  const double* tt_obs;         /* observed travel times */
  const double* weight;         /* receiver weights */
  int precision;                /* SAT_PREC_* */
  int calc_adj;                 /* 0: forward + syn only (preproc.cpp:113 is_calc_adj) */
  size_t workspace_cap;         /* bytes of per-event device workspace held at once; 0: all
                                   events when they fit in 3/4 of the free device memory,
                                   else half of it.  Above the cap a run processes consecutive
                                   chunks of the event order (forward -> adjoint -> accumulate
                                   per chunk): a pair's forward field is never kept past its
                                   adjoint.  Results are the same bits at any cap. */
} sat_batch_desc;

typedef struct sat_batch sat_batch; /* opaque: device inputs, workspace, results */

/* Upload inputs and allocate all device memory; no allocation afterwards.
 * Ownership: the per-period planes (svel, m11, m12, m22), the receiver arrays
 * and the observations may lie in host or device memory, here and in
 * sat_batch_update_media / sat_batch_update_obs (the runtime tells them
 * apart); so may the destinations of sat_batch_results and the metric and
 * outputs of sat_batch_aniso_kernels.  The axes lon_deg / lat_deg are read on
 * the host (their uniformity is checked there). */
int sat_batch_create(const sat_batch_desc* desc, sat_batch** out);
/* Replace one grid's per-period planes (next model of an inversion). */
int sat_batch_update_media(sat_batch* b, int grid, const double* svel, const double* m11,
                           const double* m12, const double* m22);
/* Replace observed times / weights (same receiver layout). */
int sat_batch_update_obs(sat_batch* b, const double* tt_obs, const double* weight);
/* Queue the whole pipeline (forward -> misfit -> adjoint -> mask -> kernel
 * accumulate -> chi) on the batch's stream.  mode: SAT_RUN_BATCHED (one
 * persistent launch), SAT_RUN_PER_SOLVE (one launch per pair, same kernels) or
 * SAT_RUN_GRAPH (the batched run captured once as a hipGraph and replayed: the
 * same work and results, one launch on the stream; set_f32_band or a change of
 * a grid's isotropy re-captures it; per-phase times are not available). */
#define SAT_RUN_BATCHED 0
#define SAT_RUN_PER_SOLVE 1
#define SAT_RUN_GRAPH 2
int sat_batch_run(sat_batch* b, int mode);
int sat_batch_sync(sat_batch* b);
/* Copy results to host.  syn[nrec_total]; adj_s: for each grid in order,
 * [nperiod][ny][nx] FP64 kernels (preproc.cpp:31 adj_s_local); chi scalar
 * (preproc.cpp:142); iters_fwd/iters_adj/flags [nevent].  Any pointer may be NULL. */
int sat_batch_results(sat_batch* b, double* syn, double* adj_s, double* chi, int* iters_fwd,
                      int* iters_adj, int* flags);
/* FP32 batches: a pair whose FP32 forward stop/continue decision lies within
 * a relative `band` of the reference tolerance (1e-4) is re-solved by the FP64
 * path in the same run, so forward iteration counts equal the reference's.
 * Default 0.02; 0 disables the re-solves. */
int sat_batch_set_f32_band(sat_batch* b, double band);
/* Per event: Linf of the last and the previous forward round (of the launch
 * that produced the result), and the number of FP64 re-solves of the last run. */
int sat_batch_conv(sat_batch* b, double* linf_last, double* linf_prev, int* n_resolved);
/* Duration in ms of the last sat_batch_run, measured with HIP events on the
 * batch's streams; solve_ms: from the start to the end of the adjoint sweeps. */
int sat_batch_last_ms(sat_batch* b, double* total_ms, double* solve_ms);
/* Per-phase ms of the last SAT_RUN_BATCHED run (HIP events on the streams the
 * kernels run on; NaN except the total after a SAT_RUN_GRAPH run), first min(n, 7) of: [0] analytic factor T0, [1] forward
 * sweeps, [2] receivers + adjoint set-up + delta, [3] adjoint sweeps, [4] FP64
 * re-solves (second stream, concurrent with [2]-[3]), [5] kernel reduction +
 * chi, [6] total.  Timing aid with no counterpart in the reference. */
int sat_batch_phase_ms(sat_batch* b, double* ms, int n);
/* Waves per pair of each grid's batched sweeps, as chosen at creation: waves[ngrid]. */
int sat_batch_waves(sat_batch* b, int* waves);
/* Scheduling options of the batched kernels (no counterpart in the reference;
 * results do not depend on them beyond FP32 rounding, and FP64 batches ignore
 * them): the autotuner's knobs besides the waves per grid.
 *   SAT_OPT_PAIR2      FP32 isotropic grids of at most 64 columns: two pairs per
 *                      wave in the forward sweeps (1), one (0), the library's
 *                      rule (-1, the default: two for one-column-per-lane grids)
 *   SAT_OPT_ADJ_COLS   FP32 one-wave grids: the adjoint on 8 columns per lane
 *                      (1, default) or on the forward's columns per lane (0)
 *   SAT_OPT_ORDER      pop each launch group's queue in the order of the
 *                      previous run's rounds, longest first (1, default), or
 *                      in batch order (0)
 *   SAT_OPT_SPECULATE  FP32 batches with a re-solve band: the pairs re-solved
 *                      in FP64 in one of the last 8 runs (and still near the
 *                      tolerance) are solved in FP64 from the start of the
 *                      run, beside the FP32 forward sweeps, which skip them
 *                      (1, default), or only the band rule's pairs, after
 *                      the FP32 forward sweeps (0).  A speculated pair's
 *                      result is its FP64 solution; one-chunk batched and
 *                      graph runs only.
 * A change re-groups the events and re-captures a graph at the next run.
 * SAT_EINVAL for an unknown option or value. */
#define SAT_OPT_PAIR2 1
#define SAT_OPT_ADJ_COLS 2
#define SAT_OPT_ORDER 3
#define SAT_OPT_SPECULATE 4
int sat_batch_set_option(sat_batch* b, int option, int value);
int sat_batch_get_option(sat_batch* b, int option, int* value);
int sat_batch_destroy(sat_batch* b);

/* ------------------------------------------------------------------------
 * Model kernels from the per-period kernels.
 * ------------------------------------------------------------------------ */

/* Replaces the isotropic part of preproc::combine_kernels (include/preproc.h,
 * src/preproc.cpp:196-342): the per-period kernels adj_s [nper][ny][nx] (as
 * sat_batch_results returns them) times the depth sensitivities sen_vs, sen_vp,
 * sen_rho [nx][ny][nz][nper] (the reference's RowMajor Tensor4r, from its
 * dispersion kernels), summed over the periods in order, into ker_* [nx][ny][nz]
 * (RowMajor Tensor3r), FP64.  mode 0: vs only (use_alpha_beta_rho false:
 * ker_vs -= att*(s_vs + s_vp*dvp/dvs + s_rho*drho/dvp*dvp/dvs) with the
 * relations of include/utils.h:198-212 at vs3d, vp3d [nx][ny][nz]); mode 1:
 * vs, vp, rho kernels; mode 2: vs, vp and rho = 0.33*vs (rho_scaling).
 * ker_vp, ker_rho (and sen_rho in mode 2, vs3d / vp3d outside mode 0) may be
 * NULL where unused.  Host pointers, synchronous; kernel_ms (may be NULL):
 * the device kernel's time.  nper <= 96.  The anisotropic (Gc, Gs) and density
 * kernels: sat_combine_kernels_ex. */
int sat_combine_kernels(int nx, int ny, int nz, int nper, int mode, const double* adj_s,
                        const double* sen_vs, const double* sen_vp, const double* sen_rho,
                        const double* vs3d, const double* vp3d, double* ker_vs, double* ker_vp,
                        double* ker_rho, double* kernel_ms);

/* preproc::combine_kernels with all of its branches (src/preproc.cpp:196-342):
 * the isotropic kernels of sat_combine_kernels, and
 *  - is_anisotropy (r1 != NULL): the vs/vp/rho terms of modes 1 and 2 scaled
 *    by 1/sqrt(1 + r1^2 + r2^2) (:253-256) and the Gc, Gs kernels (:307-338)
 *    from adj_s, adj_xi, adj_eta (sat_batch_aniso_kernels), the phase velocity
 *    svel and the sensitivities sen_gc, sen_gs;
 *  - is_kden (adj_den != NULL): the density kernel ker_den, the vs term's
 *    expression with adj_den in place of adj_s (:264-267, :294-300).
 * Layouts: adj_s, adj_xi, adj_eta, adj_den, svel [nper][ny][nx]; r1, r2
 * [nx][ny][nper] (SurfGrid r1_loc, r2_loc from prepare_aniso_media,
 * src/surf_grid.cpp:362-392); sen_* [nx][ny][nz][nper]; vs3d, vp3d and every
 * ker_* [nx][ny][nz].  Every pointer may be host or device memory.  FP64, the
 * reference's operations in its order.  nper <= 96. */
typedef struct sat_combine_desc {
  int nx, ny, nz, nper, mode;  /* mode as sat_combine_kernels */
  const double* adj_s;
  const double *sen_vs, *sen_vp, *sen_rho;
  const double *vs3d, *vp3d;   /* mode 0 */
  const double *r1, *r2;       /* NULL: isotropic */
  const double *adj_xi, *adj_eta, *svel;
  const double *sen_gc, *sen_gs;
  const double* adj_den;       /* NULL: no density kernel */
  double *ker_vs, *ker_vp, *ker_rho, *ker_gc, *ker_gs, *ker_den;
} sat_combine_desc;
int sat_combine_kernels_ex(const sat_combine_desc* d, double* kernel_ms);

/* sat_combine_kernels with the per-period kernels of a batch's last run, read
 * in place on the device (no copy through the host): grid `grid` of the
 * batch, its nx, ny, nperiod; the other arguments as sat_combine_kernels.
 * Waits for the batch's last run.  SAT_EINVAL for a batch without adjoint. */
int sat_batch_combine_kernels(sat_batch* b, int grid, int nz, int mode, const double* sen_vs, const double* sen_vp,
                              const double* sen_rho, const double* vs3d, const double* vp3d, double* ker_vs,
                              double* ker_vp, double* ker_rho, double* kernel_ms);

/* The anisotropic terms of preproc::accumulate_kernels (src/preproc.cpp:
 * 37-65, taken when is_anisotropy is set) for grid `grid` of the batch's last
 * run: per (period, node), over the period's events in batch order, with
 * Tx, Ty = gradient_2_geo(T) (include/utils.h:595-649) of the event's FP64
 * travel time and Ta its masked adjoint field,
 *   adj_xi  += Ta * (Tx^2 (c^2 - a^2) + 2 TxTy (ac - bc) + Ty^2 (b^2 - c^2))
 *   adj_eta += Ta * (2 ac Tx^2 - 2 (ab + c^2) TxTy + 2 bc Ty^2).
 * a, b, c: SurfGrid's a, b, c (the topography metric: m11, m22 and -m12 of
 * sat_topo_metric), [nperiod][ny][nx] host arrays; adj_xi, adj_eta the same
 * shape.  Waits for the last run (any mode; FP32 pairs re-solved in FP64 use
 * their FP64 fields); synchronous.  SAT_EINVAL before the first run, for a
 * batch without adjoint, or for one whose runs are split into chunks (its
 * fields are not all resident at the end of a run). */
int sat_batch_aniso_kernels(sat_batch* b, int grid, const double* a, const double* bm, const double* c,
                            double* adj_xi, double* adj_eta);

/* The same terms formed inside every run, chunked runs included (each chunk
 * adds its events before its workspace is reused): register grid `grid`'s
 * metric a, bm, c ([nperiod][ny][nx], host or device) once; from the next
 * run on, sat_batch_aniso_kernels(b, grid, NULL, NULL, NULL, adj_xi, adj_eta)
 * returns what the last run formed.  All three NULL unregisters.  The
 * counterpart of is_anisotropy in accumulate_kernels (src/preproc.cpp:37-65). */
int sat_batch_set_aniso_metric(sat_batch* b, int grid, const double* a, const double* bm, const double* c);

/* The density kernel of forward_for_event (is_kden, src/preproc.cpp:127-136,
 * accumulate_kernels :32-34): with on != 0 every run also solves, for every
 * pair, the adjoint of the receiver sources -weight on the pair's forward
 * field, and accumulates kden_s[iper] += field * svel^-3 in FP64 in batch
 * order.  The field is not masked: the reference masks adj_field a second
 * time (:136), not kden_field.  The run's adjoint work doubles.
 * sat_batch_kden_results copies the last run's kden_s (the layout of adj_s in
 * sat_batch_results; host or device memory); SAT_EINVAL when that run did not
 * form it. */
int sat_batch_set_kden(sat_batch* b, int on);
int sat_batch_kden_results(sat_batch* b, double* kden_s);

/* The anisotropic metric of a topographic surface for every period:
 * SurfGrid::build_media_matrix_with_topo (src/surf_grid.cpp:109-145, :184-190)
 * -- the surface z [ny][nx] (km, ix fastest: the reference's column-major
 * (nx, ny) topo.z) smoothed per period with gaussian_smooth_geo_2
 * (include/utils.h:505-573) at sigma_deg[iper] degrees (<= 0: not smoothed),
 * its geographic gradient fx, fy (gradient_2_geo, include/utils.h:595-649),
 * then m11 = (1+fy^2)/D, m22 = (1+fx^2)/D, m12 = -fx*fy/D with
 * D = 1+fx^2+fy^2, written [nper][ny][nx] (the batch's period-major planes).
 * zs_out (may be NULL): the smoothed surfaces [nper][ny][nx].  The reference's
 * sigma is v(T)*T*wavelen_factor/(2 pi) km / (R_EARTH*DEG2RAD) (:115-118).
 * Axes uniform and ascending (dlon = lon[1]-lon[0], dlat likewise, as the
 * reference reads them); 3 <= nx <= 2048, ny >= 3, nper <= 65535, and
 * 2*int(3 sigma/dlon)+1 <= 4095.  Host pointers, synchronous; kernel_ms (may
 * be NULL): the device time of the smoothing and metric kernels. */
int sat_topo_metric(int nx, int ny, int nper, const double* lon_deg, const double* lat_deg, const double* z,
                    const double* sigma_deg, double* m11, double* m12, double* m22, double* zs_out,
                    double* kernel_ms);

/* Surface-wave dispersion of ncol layered columns: the reference's
 * surfdisp96_c (src/surfker/surfdisp96_c_api.f90:5, src/surfker/surfdisp96.f:57)
 * batched over columns.  thk, vp, vs, rho: [ncol][nlayer] REAL*4, the last
 * layer the half-space; t: [kmax] periods (s); iflsph 0 flat / 1 spherical
 * earth; iwave 1 Love / 2 Rayleigh; mode >= 1 (1 fundamental); igr 0 phase /
 * > 0 group velocity.  cg: [ncol][kmax] velocities (float-rounded, as the
 * reference returns them).  flags (may be NULL): [ncol], 1 where the
 * fundamental mode has no root (the reference stops the program there; cg is
 * 0 from that period on).  2 <= nlayer <= 512, 1 <= kmax <= 1024.  Host
 * pointers, synchronous; kernel_ms (may be NULL): the device time. */
int sat_surfdisp96(int ncol, int nlayer, const float* thk, const float* vp, const float* vs, const float* rho,
                   int iflsph, int iwave, int mode, int igr, int kmax, const double* t, double* cg, int* flags,
                   double* kernel_ms);

/* Phase- or group-velocity maps of a 3-D model: SurfGrid::fwdsurf
 * (src/surf_grid.cpp:194-245) -- per column (ix, iy) the model at the depth
 * nodes zgrids[nz] (vs3d, and vp3d/rho3d or, when both are NULL, Brocher's
 * vp(vs) and rho(vp), include/utils.h:185-187,223-225; all [nx][ny][nz] FP64,
 * the reference's RowMajor Tensor3r), layered as surfker::build_disp_req does
 * (src/surfdisp.cpp:60-111: nz + 1 layers, the last node's layer repeated and
 * a half-space), solved by surfdisp96 at nper periods.  svel:
 * [nper][ny][nx] (the batch's period-major planes).  flags as for
 * sat_surfdisp96, [ny][nx].  nz >= 2.  The reference calls it with iflsph 1,
 * mode 1 (include/config.h:79-80). */
int sat_fwdsurf(int nx, int ny, int nz, const double* zgrids, const double* vs3d, const double* vp3d,
                const double* rho3d, int nper, const double* periods, int iflsph, int iwave, int mode, int igr,
                double* svel, int* flags, double* kernel_ms);

/* Library identity: device name, build arch; returns SAT_EDEVICE without a GPU. */
int sat_device_info(char* buf, int buflen);

#ifdef __cplusplus
}
#endif
#endif
