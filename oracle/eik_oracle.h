/* eik_oracle.h -- CPU FP64 restatement of SurfATTPP's eikonal hot path.
 * TEST INFRASTRUCTURE ONLY (the checker); see eik_oracle.c. */
#ifndef SAT_EIK_ORACLE_H
#define SAT_EIK_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define SAT_ORACLE_OK 0
#define SAT_ORACLE_EINVAL 22
#define SAT_ORACLE_ENOMEM 12

#define SAT_FLAG_NOT_CONVERGED 1
#define SAT_FLAG_SRC_ON_BOUNDARY 2

int sat_oracle_forward(int nx, int ny, const double* xx_deg, const double* yy_deg,
                       const double* spha, const double* sphb, const double* sphc,
                       const double* fun, double x0_deg, double y0_deg,
                       int max_iter, double* T_out, int* iters_out, int* flags_out);

int sat_oracle_adjoint(int nx, int ny, const double* xx_deg, const double* yy_deg,
                       const double* spha, const double* sphb, const double* sphc,
                       const double* T, int nr, const double* xrec_deg, const double* yrec_deg,
                       const double* sourceAdj, int max_iter, double* Ta_out, int* iters_out);

void sat_oracle_mask(int nx, int ny, const double* xx, const double* yy, double* Ta,
                     double x0, double y0);

void sat_oracle_interp2d(int nx, int ny, const double* xgrid, const double* ygrid,
                         const double* z, int nq, const double* xq, const double* yq,
                         double* zq);

int sat_oracle_forward_for_event(int nx, int ny, const double* xgrids, const double* ygrids,
                                 int nperiod, const double* svel, const double* m11,
                                 const double* m12, const double* m22,
                                 int nevent, const int* ev_iper, const double* ev_lon,
                                 const double* ev_lat, const int* ev_rec_off, const int* ev_nrec,
                                 const double* rec_lon, const double* rec_lat,
                                 const double* tt_obs, const double* weight,
                                 int is_calc_adj, double* syn, double* adj_s, double* chi_out,
                                 int* iters_fwd, int* iters_adj, int* flags);

int sat_oracle_forward_for_event_kden(int nx, int ny, const double* xgrids, const double* ygrids,
                                      int nperiod, const double* svel, const double* m11,
                                      const double* m12, const double* m22,
                                      int nevent, const int* ev_iper, const double* ev_lon,
                                      const double* ev_lat, const int* ev_rec_off, const int* ev_nrec,
                                      const double* rec_lon, const double* rec_lat,
                                      const double* tt_obs, const double* weight,
                                      int is_calc_adj, double* syn, double* adj_s, double* chi_out,
                                      int* iters_fwd, int* iters_adj, int* flags, double* kden_s);

int sat_oracle_combine_kernels(int nx, int ny, int nz, int nper, int mode, const double* adj_tt,
                               const double* sen_vs, const double* sen_vp, const double* sen_rho,
                               const double* vs3d, const double* vp3d, double* ker_vs, double* ker_vp,
                               double* ker_rho);

int sat_oracle_combine_kernels_ex(int nx, int ny, int nz, int nper, int mode, const double* adj_tt,
                                  const double* sen_vs, const double* sen_vp, const double* sen_rho,
                                  const double* vs3d, const double* vp3d, const double* r1, const double* r2,
                                  const double* adj_xi, const double* adj_eta, const double* svel,
                                  const double* sen_gc, const double* sen_gs, const double* adj_den,
                                  double* ker_vs, double* ker_vp, double* ker_rho, double* ker_gc,
                                  double* ker_gs, double* ker_den);

#ifdef __cplusplus
}
#endif
#endif
