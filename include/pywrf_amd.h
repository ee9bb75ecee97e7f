/*
 * pywrf_amd: C ABI of the MI355X (gfx950) kernels for the ARW RK3 step of
 * PyWRF (the body of WrfSolver.solve's time loop, pywrf/solver.py:454-1910).
 *
 * One entry point per reference function on the path.  Arguments keep the
 * reference's order with tensors as plain device pointers (FP32, contiguous,
 * (nz, ny, nx) with i fastest; 2-D (ny, nx); level arrays (nz)), Python
 * floats as double, integer flags as int, staggering strings as one char,
 * and the WRF index sets packed in pywrf_grid_t.  Every call is asynchronous
 * on `stream` (a hipStream_t) and returns 0 or an error code: a hipError_t
 * from the launch, PYWRF_E_ARG for an argument the reference rejects (it
 * raises NameError for an unknown staggering flag) or PYWRF_E_RANGE for index
 * sets whose stencils would leave the allocation.
 *
 * Where the reference returns a fresh tensor (clone / torch.where / new
 * arithmetic result) the caller passes the output buffer explicitly
 * (`*_out`); where it mutates an argument in place the same pointer is both
 * read and written.
 */
#ifndef PYWRF_AMD_H
#define PYWRF_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef void* pywrf_stream_t; /* hipStream_t */

#define PYWRF_E_ARG (-1)
#define PYWRF_E_RANGE (-2)

/* Allocation shape and the WRF index sets (config_params.py:67-99).  Functions
 * whose reference signature has no patch indices ignore ips..kpe. */
typedef struct pywrf_grid {
  int nz, ny, nx;
  int ids, ide, jds, jde, kds, kde;
  int ims, ime, jms, jme, kms, kme;
  int ips, ipe, jps, jpe, kps, kpe;
  int its, ite, jts, jte, kts, kte;
} pywrf_grid_t;

/* Library version and the value of a named physical constant (for host-side
 * consistency checks); returns 0 when the name is known. */
int pywrf_version(void);
int pywrf_constant(const char* name, double* value);

/* ---- RK-stage preparation: wrf_dynamics.py:30-159 ---------------------- */
/* calculate_full, wrf_dynamics.py:30-38 */
int pywrf_calculate_full(const pywrf_grid_t* g, float* mut, const float* mub, const float* mu, pywrf_stream_t s);
/* calc_mu_uv, wrf_dynamics.py:41-61 */
int pywrf_calc_mu_uv(const pywrf_grid_t* g, const float* mu, const float* mub, float* muu, float* muv,
                     pywrf_stream_t s);
/* couple_momentum, wrf_dynamics.py:64-86 */
int pywrf_couple_momentum(const pywrf_grid_t* g, const float* muu, float* ru, const float* u, const float* msfu,
                          const float* muv, float* rv, const float* v, const float* msfv, const float* msfvx_inv,
                          const float* mut, float* rw, const float* w, const float* msft, pywrf_stream_t s);
/* calc_ww_cp, wrf_dynamics.py:89-118 */
int pywrf_calc_ww_cp(const pywrf_grid_t* g, const float* u, const float* v, const float* muu, const float* muv,
                     const float* c1h, float* ww, double rdx, double rdy, const float* msftx,
                     const float* msfuy, const float* msfvx_inv, const float* dnw, pywrf_stream_t s);
/* calc_cq, wrf_dynamics.py:121-136 (moist: n_moist consecutive 3-D fields) */
int pywrf_calc_cq(const pywrf_grid_t* g, const float* moist, int n_moist, float* cqu, float* cqv, float* cqw,
                  pywrf_stream_t s);
/* calc_alt, wrf_dynamics.py:139-147 */
int pywrf_calc_alt(const pywrf_grid_t* g, float* alt, const float* al, const float* alb, pywrf_stream_t s);
/* calc_php, wrf_dynamics.py:150-159 */
int pywrf_calc_php(const pywrf_grid_t* g, float* php, const float* ph, const float* phb, pywrf_stream_t s);

/* ---- open lateral boundaries: wrf_dynamics.py:164-278 ------------------- */
/* set_physical_bc2d, wrf_dynamics.py:164-212 (variable: u v w t x y r p, any case) */
int pywrf_set_physical_bc2d(const pywrf_grid_t* g, float* dat, char variable, pywrf_stream_t s);
/* set_physical_bc3d, wrf_dynamics.py:216-278 (variable: u v w t d e x y f r p, U V) */
int pywrf_set_physical_bc3d(const pywrf_grid_t* g, float* dat, char variable, pywrf_stream_t s);

/* A group of n consecutive set_physical_bc3d / set_physical_bc2d calls on
 * distinct fields (as the reference driver issues them, e.g. solver.py:781-843)
 * in one launch: fields[f] is a device pointer, variables[f] its flag,
 * is2d[f] selects the 2-D form (index sets of g2) over the 3-D form (g3).
 * No counterpart in the reference: a batching of the two calls above. */
int pywrf_set_physical_bc_group(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n, void* const* fields,
                                const char* variables, const int* is2d, pywrf_stream_t s);

/* Zero n distinct device arrays of counts[f] floats each in one launch: the
 * runs of tensor.zero_() calls of the reference driver (solver.py:848-851,
 * 975-981).  No counterpart in the reference API: a batching of zero_(). */
int pywrf_zero_group(int n, void* const* fields, const long long* counts, pywrf_stream_t s);

/* ---- first-stage physics prep and Smagorinsky: wrf_dynamics.py:314-1076 - */
/* phy_prep, wrf_dynamics.py:314-411 (pi_phy_out, t_phy_out: fresh arrays) */
int pywrf_phy_prep(const pywrf_grid_t* g, const float* mut, const float* u, const float* v, const float* p,
                   const float* pb, const float* alt, const float* ph, const float* phb, const float* t,
                   const float* moist, int n_moist, float* rho, float* th_phy, float* p_phy, float* u_phy,
                   float* v_phy, float* p8w, float* t8w, float* z, float* z_at_w, float* dz8w, float* p_hyd,
                   float* p_hyd_w, const float* dnw, const float* fzm, const float* fzp, double p_top,
                   float* pi_phy_out, float* t_phy_out, pywrf_stream_t s);
/* calculate_phy_tend, wrf_dynamics.py:449-488 (scalar_tend: 3 fields) */
int pywrf_calculate_phy_tend(const pywrf_grid_t* g, const float* mut, float* rublten, float* rvblten,
                             float* rthblten, float* rqvblten, float* rqcblten, float* rqiblten,
                             float* scalar_tend, pywrf_stream_t s);
/* compute_diff_metrics, wrf_dynamics.py:491-536 */
int pywrf_compute_diff_metrics(const pywrf_grid_t* g, const float* ph, const float* phb, float* z, float* rdz,
                               float* rdzw, float* zx, float* zy, double rdx, double rdy, pywrf_stream_t s);
/* cal_deform_and_div, wrf_dynamics.py:539-813 (cf1..cf3: 0-d arrays) */
int pywrf_cal_deform_and_div(const pywrf_grid_t* g, const float* u, const float* v, const float* w, float* div,
                             float* defor11, float* defor22, float* defor33, float* defor12, float* defor13,
                             float* defor23, const float* msfux, const float* msfuy, const float* msfvx,
                             const float* msfvy, const float* msftx, const float* msfty, double rdx, double rdy,
                             const float* dn, const float* dnw, const float* rdz, const float* rdzw,
                             const float* fnm, const float* fnp, const float* cf, const float* zx,
                             const float* zy, pywrf_stream_t s);
/* calculate_km_kh, wrf_dynamics.py:816-995 (BN2_out, xkmh_out: fresh arrays;
 * xkmh is also updated in place on the interior as the reference does) */
int pywrf_calculate_km_kh(const pywrf_grid_t* g, float* xkmh, float* xkmv, float* xkhh, float* xkhv,
                          const float* defor11, const float* defor22, const float* defor12, const float* p8w,
                          const float* t8w, const float* theta, const float* t, const float* p,
                          const float* moist, double dx, double dy, const float* rdz, const float* rdzw,
                          const float* cf, const float* msftx, const float* msfty, float* BN2_out,
                          float* xkmh_out, pywrf_stream_t s);
/* update_phy_ten with add_a2a / add_a2c_u / add_a2c_v, wrf_dynamics.py:998-1076 */
int pywrf_update_phy_ten(const pywrf_grid_t* g, float* rt_tendf, float* ru_tendf, float* rv_tendf,
                         float* moist_tendf, const float* rthblten, const float* rublten, const float* rvblten,
                         const float* rqvblten, const float* rqcblten, const float* rqiblten, pywrf_stream_t s);

/* ---- large-step tendencies: wrf_dynamics.py:1112-3148 ------------------- */
/* advect_u, wrf_dynamics.py:1112-1307 (u_old, c1, c2, mut, time_step unused as in the reference) */
int pywrf_advect_u(const pywrf_grid_t* g, const float* u, float* tendency, const float* ru, const float* rv,
                   const float* rom, const float* msfux, const float* fzm, const float* fzp, double rdx, double rdy,
                   const float* rdzw, pywrf_stream_t s);
/* advect_v, wrf_dynamics.py:1332-1532 */
int pywrf_advect_v(const pywrf_grid_t* g, const float* v, float* tendency, const float* ru, const float* rv,
                   const float* rom, const float* msfvx, const float* msfvy, const float* fzm, const float* fzp,
                   double rdx, double rdy, const float* rdzw, pywrf_stream_t s);
/* advect_w, wrf_dynamics.py:1535-1822 */
int pywrf_advect_w(const pywrf_grid_t* g, const float* w, float* tendency, const float* ru, const float* rv,
                   const float* rom, const float* msftx, const float* fzm, const float* fzp, double rdx, double rdy,
                   const float* rdzu, pywrf_stream_t s);
/* advect_scalar, wrf_dynamics.py:1841-2028 */
int pywrf_advect_scalar(const pywrf_grid_t* g, const float* field, float* tendency, const float* ru,
                        const float* rv, const float* rom, const float* msftx, const float* fzm, const float* fzp,
                        double rdx, double rdy, const float* rdzw, pywrf_stream_t s);
/* rhs_ph, wrf_dynamics.py:2031-2297 (cfn, cfn1: 0-d arrays) */
int pywrf_rhs_ph(const pywrf_grid_t* g, float* ph_tend, const float* u, const float* v, const float* ww,
                 const float* ph, const float* phb, const float* w, const float* mut, const float* muuf,
                 const float* muvf, const float* fnm, const float* fnp, const float* rdnw, const float* cfn,
                 const float* cfn1, double rdx, double rdy, const float* msfux, const float* msfvy,
                 const float* msfty, pywrf_stream_t s);
/* The five calls above into a tendency array the caller would have zeroed
 * first (solver.py:971-975 zero ru, rv, rw, t and ph_tend before the calls
 * of :987-1036): the whole array is written, 0 + the tendency inside the
 * call's index box and 0 elsewhere, and the array is not read.  Same
 * arguments and results as zeroing and then making the call.  No counterpart
 * in the reference: the zeroing folded into the first writer. */
int pywrf_advect_u_fresh(const pywrf_grid_t* g, const float* u, float* tendency, const float* ru, const float* rv,
                         const float* rom, const float* msfux, const float* fzm, const float* fzp, double rdx,
                         double rdy, const float* rdzw, pywrf_stream_t s);
int pywrf_advect_v_fresh(const pywrf_grid_t* g, const float* v, float* tendency, const float* ru, const float* rv,
                         const float* rom, const float* msfvx, const float* msfvy, const float* fzm,
                         const float* fzp, double rdx, double rdy, const float* rdzw, pywrf_stream_t s);
int pywrf_advect_w_fresh(const pywrf_grid_t* g, const float* w, float* tendency, const float* ru, const float* rv,
                         const float* rom, const float* msftx, const float* fzm, const float* fzp, double rdx,
                         double rdy, const float* rdzu, pywrf_stream_t s);
int pywrf_advect_scalar_fresh(const pywrf_grid_t* g, const float* field, float* tendency, const float* ru,
                              const float* rv, const float* rom, const float* msftx, const float* fzm,
                              const float* fzp, double rdx, double rdy, const float* rdzw, pywrf_stream_t s);
int pywrf_rhs_ph_fresh(const pywrf_grid_t* g, float* ph_tend, const float* u, const float* v, const float* ww,
                       const float* ph, const float* phb, const float* w, const float* mut, const float* muuf,
                       const float* muvf, const float* fnm, const float* fnp, const float* rdnw, const float* cfn,
                       const float* cfn1, double rdx, double rdy, const float* msfux, const float* msfvy,
                       const float* msfty, pywrf_stream_t s);
/* horizontal_pressure_gradient, wrf_dynamics.py:2300-2421 (cf: cf1, cf2, cf3 as 3 floats) */
int pywrf_horizontal_pressure_gradient(const pywrf_grid_t* g, float* ru_tend, float* rv_tend, const float* ph,
                                       const float* alt, const float* p, const float* pb, const float* al,
                                       const float* php, const float* cqu, const float* cqv, const float* muu,
                                       const float* muv, const float* mu, const float* fnm, const float* fnp,
                                       const float* rdnw, const float* cf, double rdx, double rdy,
                                       const float* msfux, const float* msfuy, const float* msfvx,
                                       const float* msfvy, pywrf_stream_t s);
/* pg_buoy_w, wrf_dynamics.py:2424-2460 (cqw overwritten with 1/(1+cqw) as in the reference) */
int pywrf_pg_buoy_w(const pywrf_grid_t* g, float* rw_tend, const float* p, float* cqw, const float* muf,
                    const float* mubf, const float* rdnw, const float* rdn, double g_, const float* msfty,
                    pywrf_stream_t s);
/* coriolis, wrf_dynamics.py:2463-2532 */
int pywrf_coriolis(const pywrf_grid_t* g, const float* ru, const float* rv, const float* rw, float* ru_tend,
                   float* rv_tend, float* rw_tend, const float* msftx, const float* msfty, const float* msfux,
                   const float* msfuy, const float* msfvx, const float* msfvy, const float* f, const float* e,
                   const float* sina, const float* cosa, const float* fzm, const float* fzp, pywrf_stream_t s);
/* curvature, wrf_dynamics.py:2535-2620 */
int pywrf_curvature(const pywrf_grid_t* g, const float* ru, const float* rv, const float* rw, const float* u,
                    const float* v, float* ru_tend, float* rv_tend, float* rw_tend, const float* msfuy,
                    const float* msfvx, const float* msfvy, const float* msftx, const float* msfty,
                    const float* fzm, const float* fzp, double rdx, double rdy, pywrf_stream_t s);
/* coriolis then curvature, as the driver calls them back to back
 * (solver.py:1063-1079), in one pass over the tendencies: the arguments of
 * both (ru, rv, rw, u, v, the three tendencies updated in place, the map
 * factors, f, e, sina, cosa, fzm, fzp, rdx, rdy).  Equal to the two calls.
 * No counterpart in the reference: a fusion. */
int pywrf_coriolis_curvature(const pywrf_grid_t* g, const float* ru, const float* rv, const float* rw,
                             const float* u, const float* v, float* ru_tend, float* rv_tend, float* rw_tend,
                             const float* msftx, const float* msfty, const float* msfux, const float* msfuy,
                             const float* msfvx, const float* msfvy, const float* f, const float* e,
                             const float* sina, const float* cosa, const float* fzm, const float* fzp, double rdx,
                             double rdy, pywrf_stream_t s);
/* horizontal_pressure_gradient (ru_tend, rv_tend), pg_buoy_w (rw_tend, cqw)
 * and coriolis_curvature (the three tendencies) of one RK stage
 * (solver.py:1038-1080) in one launch, each point in that order: same results
 * as the three calls.  The buoyancy's mu is the gradient's mu, as in the
 * reference's calls; fnm, fnp serve as fzm, fzp.  No counterpart in the
 * reference: a fusion. */
int pywrf_pgf_buoy_coriolis_curvature(const pywrf_grid_t* g, float* ru_tend, float* rv_tend, float* rw_tend,
                                      const float* ph, const float* alt, const float* p, const float* pb,
                                      const float* al, const float* php, const float* cqu, const float* cqv,
                                      float* cqw, const float* muu, const float* muv, const float* mu,
                                      const float* mub, const float* fnm, const float* fnp, const float* rdnw,
                                      const float* rdn, const float* cf, double g_, const float* ru,
                                      const float* rv, const float* rw, const float* u, const float* v,
                                      const float* msftx, const float* msfty, const float* msfux,
                                      const float* msfuy, const float* msfvx, const float* msfvy, const float* f,
                                      const float* e, const float* sina, const float* cosa, double rdx,
                                      double rdy, pywrf_stream_t s);
/* horizontal_diffusion, wrf_dynamics.py:2623-2838 (name: 'u', 'v', 'w' or a scalar name) */
int pywrf_horizontal_diffusion(const pywrf_grid_t* g, char name, const float* field, float* tendency,
                               const float* mut, const float* msfux, const float* msfuy, const float* msfvx,
                               const float* msfvx_inv, const float* msfvy, const float* msftx, const float* msfty,
                               const float* xkmhd, double rdx, double rdy, pywrf_stream_t s);
/* horizontal_diffusion_3dmp, wrf_dynamics.py:2841-2915 */
int pywrf_horizontal_diffusion_3dmp(const pywrf_grid_t* g, const float* field, float* tendency,
                                    const float* mut, const float* base_3d, const float* msfux,
                                    const float* msfuy, const float* msfvx, const float* msfvx_inv,
                                    const float* msfvy, const float* msftx, const float* msfty,
                                    const float* xkmhd, double rdx, double rdy, pywrf_stream_t s);
/* vertical_diffusion, wrf_dynamics.py:2917-2959 (name 'w' or 'm'; others leave tendency untouched) */
int pywrf_vertical_diffusion(const pywrf_grid_t* g, char name, const float* field, float* tendency,
                             const float* alt, const float* mut, const float* rdn, const float* rdnw, double kvdif,
                             pywrf_stream_t s);
/* vertical_diffusion_u, wrf_dynamics.py:2962-2991 */
int pywrf_vertical_diffusion_u(const pywrf_grid_t* g, const float* field, float* tendency, const float* u_base,
                               const float* alt, const float* muu, const float* rdn, const float* rdnw,
                               double kvdif, pywrf_stream_t s);
/* vertical_diffusion_v, wrf_dynamics.py:2994-3024 */
int pywrf_vertical_diffusion_v(const pywrf_grid_t* g, const float* field, float* tendency, const float* v_base,
                               const float* alt, const float* muv, const float* rdn, const float* rdnw,
                               double kvdif, pywrf_stream_t s);
/* vertical_diffusion_mp, wrf_dynamics.py:3026-3055 */
int pywrf_vertical_diffusion_mp(const pywrf_grid_t* g, const float* field, float* tendency, const float* base,
                                const float* alt, const float* mut, const float* rdn, const float* rdnw,
                                double kvdif, pywrf_stream_t s);
/* vertical_diffusion_3dmp, wrf_dynamics.py:3058-3086 */
int pywrf_vertical_diffusion_3dmp(const pywrf_grid_t* g, const float* field, float* tendency,
                                  const float* base_3d, const float* alt, const float* mut, const float* rdn,
                                  const float* rdnw, double kvdif, pywrf_stream_t s);
/* The eight diffusion calls of the first stage (solver.py:1079-1141):
 * horizontal_diffusion 'u', 'v', 'w' (coefficient xkmh), horizontal_diffusion_3dmp
 * of t about t_init (xkhh), vertical_diffusion_u, _v, vertical_diffusion 'w'
 * (mu: mut) with kvdif_m and vertical_diffusion_3dmp of t about t_init (mut)
 * with kvdif_t, in one launch with the same result as the calls in that
 * order.  The four tendencies are distinct arrays that no input aliases.  No
 * counterpart in the reference: a fusion. */
int pywrf_diffusion_dry(const pywrf_grid_t* g, const float* u, float* ru_tendf, const float* v, float* rv_tendf,
                        const float* w, float* rw_tendf, const float* t, float* t_tendf, const float* t_init,
                        const float* mut, const float* msfux, const float* msfuy, const float* msfvx,
                        const float* msfvx_inv, const float* msfvy, const float* msftx, const float* msfty,
                        const float* xkmh, const float* xkhh, const float* u_base, const float* v_base,
                        const float* alt, const float* muu, const float* muv, const float* rdn, const float* rdnw,
                        double kvdif_m, double kvdif_t, double rdx, double rdy, pywrf_stream_t s);
/* rk_addtend_dry, wrf_dynamics.py:3100-3148 */
int pywrf_rk_addtend_dry(const pywrf_grid_t* g, float* ru_tend, float* rv_tend, float* rw_tend, float* ph_tend,
                         float* t_tend, float* ru_tendf, float* rv_tendf, float* rw_tendf, float* ph_tendf,
                         float* t_tendf, const float* u_save, const float* v_save, const float* w_save,
                         const float* ph_save, const float* t_save, float* mu_tend, const float* mu_tendf,
                         int rk_step, const float* h_diabatic, const float* mut, const float* msfty,
                         const float* msfuy, const float* msfvx, const float* msfvx_inv, pywrf_stream_t s);

/* ---- acoustic small step: wrf_dynamics.py:3152-3930 -------------------- */
/* small_step_prep, wrf_dynamics.py:3152-3266 (u_out..ph_out: the five fresh perturbation arrays) */
int pywrf_small_step_prep(const pywrf_grid_t* g, float* u_1, const float* u_2, float* v_1, const float* v_2,
                          float* w_1, const float* w_2, float* t_1, const float* t_2, float* ph_1,
                          const float* ph_2, const float* mub, float* mu_1, float* mu_2, const float* muu,
                          float* muus, const float* muv, float* muvs, const float* mut, float* muts, float* mudf,
                          float* u_save, float* v_save, float* w_save, float* t_save, float* ph_save,
                          float* mu_save, const float* ww, float* ww_save, float* c2a, const float* pb,
                          const float* p, const float* alt, const float* msfuy, const float* msfvx_inv,
                          const float* msfty, int rk_step, float* u_out, float* v_out, float* w_out, float* t_out,
                          float* ph_out, pywrf_stream_t s);
/* calc_p_rho, wrf_dynamics.py:3269-3313 (al_out, p_out fresh; pm1 in place, or the fresh copy at step 0) */
int pywrf_calc_p_rho(const pywrf_grid_t* g, const float* al_in, const float* p_in, const float* ph,
                     const float* alt, const float* t_2, const float* t_1, const float* c2a, float* pm1,
                     const float* mu, const float* mut, double t0, const float* rdnw, double smdiv, int step,
                     float* al_out, float* p_out, pywrf_stream_t s);
/* calc_coef_w, wrf_dynamics.py:3316-3388 */
int pywrf_calc_coef_w(const pywrf_grid_t* g, float* a, float* alpha, float* gamma, const float* mut,
                      const float* cqw, const float* rdn, const float* rdnw, const float* c2a, double dts,
                      double g_, double epssm, int top_lid, pywrf_stream_t s);
/* advance_uv, wrf_dynamics.py:3391-3525 (u_out, v_out fresh) */
int pywrf_advance_uv(const pywrf_grid_t* g, const float* u_in, const float* ru_tend, const float* v_in,
                     const float* rv_tend, const float* p, const float* pb, const float* ph, const float* php,
                     const float* alt, const float* al, const float* mu, const float* muu, const float* cqu,
                     const float* muv, const float* cqv, const float* mudf, const float* msfux, const float* msfuy,
                     const float* msfvx, const float* msfvx_inv, const float* msfvy, double rdx, double rdy,
                     double dts, const float* cf, const float* fnm, const float* fnp, double emdiv,
                     const float* rdnw, int spec_zone, float* u_out, float* v_out, pywrf_stream_t s);
/* advance_mu_t, wrf_dynamics.py:3528-3638 (the seven *_out arrays fresh) */
int pywrf_advance_mu_t(const pywrf_grid_t* g, const float* ww_in, const float* ww_1, const float* u,
                       const float* u_1, const float* v, const float* v_1, const float* mu_in, const float* mut,
                       const float* muave_in, const float* muts_in, const float* muu, const float* muv,
                       const float* mudf_in, const float* t_in, const float* t_1, const float* t_ave_in,
                       const float* ft, const float* mu_tend, double rdx, double rdy, double dts, double epssm,
                       const float* dnw, const float* fnm, const float* fnp, const float* rdnw,
                       const float* msfuy, const float* msfvx_inv, const float* msftx, const float* msfty,
                       float* ww_out, float* t_out, float* t_ave_out, float* muave_out, float* muts_out,
                       float* mudf_out, float* mu_out, pywrf_stream_t s);
/* advance_w, wrf_dynamics.py:3641-3796 (t_2ave_out, w_out, ph_out fresh) */
int pywrf_advance_w(const pywrf_grid_t* g, const float* w_in, const float* rw_tend, const float* ww,
                    const float* w_save, const float* u, const float* v, const float* mut, const float* muave,
                    const float* muts, const float* t_2ave_in, const float* t_2, const float* t_1,
                    const float* ph_in, const float* ph_1, const float* phb, const float* ph_tend,
                    const float* ht, const float* c2a, const float* cqw, const float* alt, const float* a,
                    const float* alpha, const float* gamma, double rdx, double rdy, double dts, double t0,
                    double epssm, const float* fnm, const float* fnp, const float* rdnw, const float* rdn,
                    const float* cf, const float* msftx, const float* msfty, float* t_2ave_out, float* w_out,
                    float* ph_out, pywrf_stream_t s);
/* sumflux, wrf_dynamics.py:3799-3839 */
int pywrf_sumflux(const pywrf_grid_t* g, const float* ru, const float* rv, const float* ww, const float* u_lin,
                  const float* v_lin, const float* ww_lin, const float* muu, const float* muv, float* ru_m,
                  float* rv_m, float* ww_m, const float* msfuy, const float* msfvx_inv, int iteration,
                  int number_of_small_timesteps, pywrf_stream_t s);
/* calc_mu_uv_1, wrf_dynamics.py:3844-3860 */
int pywrf_calc_mu_uv_1(const pywrf_grid_t* g, const float* mu, float* muu, float* muv, pywrf_stream_t s);
/* small_step_finish, wrf_dynamics.py:3863-3930 (u..t_out fresh; ww, mu_2 in place) */
int pywrf_small_step_finish(const pywrf_grid_t* g, const float* u_2, const float* v_2, const float* w_2,
                            const float* t_2, const float* ph_2, float* ww, const float* ww1, float* mu_2,
                            const float* mut, const float* muts, const float* muu, const float* muus,
                            const float* muv, const float* muvs, const float* u_save, const float* v_save,
                            const float* w_save, const float* t_save, const float* ph_save, const float* mu_save,
                            const float* msfuy, const float* msfvx, const float* msfty, const float* h_diabatic,
                            int number_of_small_timesteps, double dts, int rk_step, int rk_order, float* u_out,
                            float* v_out, float* w_out, float* ph_out, float* t_out, pywrf_stream_t s);

/* ---- scalar transport, EOS, surface w: wrf_dynamics.py:3982-4324 ------- */
/* rk_scalar_tend, wrf_dynamics.py:3982-4051 (scalar, scalar_tends, advect_tend: (n, nz, ny, nx)) */
int pywrf_rk_scalar_tend(const pywrf_grid_t* g, int scs, int sce, int rk_step, const float* ru, const float* rv,
                         const float* ww, const float* mut, const float* alt, const float* scalar,
                         float* scalar_tends, float* advect_tend, float* h_tendency, float* z_tendency,
                         const float* base, int moist_step, const float* fnm, const float* fnp,
                         const float* msfux, const float* msfuy, const float* msfvx, const float* msfvx_inv,
                         const float* msfvy, const float* msftx, const float* msfty, double rdx, double rdy,
                         const float* rdn, const float* rdnw, double kvdif, const float* xkmhd,
                         pywrf_stream_t s);
/* Two rk_scalar_tend calls that share everything but the species array, its
 * range, its two tendency arrays and the moist flag (the moist and the
 * tracer call of solver.py:1548-1610), in the same launches; h_tendency and
 * z_tendency are zeroed once.  Same results as the two calls.  No
 * counterpart in the reference: a batching. */
int pywrf_rk_scalar_tend_pair(const pywrf_grid_t* g, int rk_step, int scs, int sce, const float* scalar,
                              float* scalar_tends, float* advect_tend, int moist_step, int scs2, int sce2,
                              const float* scalar2, float* scalar_tends2, float* advect_tend2, int moist_step2,
                              const float* ru, const float* rv, const float* ww, const float* mut,
                              const float* alt, float* h_tendency, float* z_tendency, const float* base,
                              const float* fnm, const float* fnp, const float* msfux, const float* msfuy,
                              const float* msfvx, const float* msfvx_inv, const float* msfvy, const float* msftx,
                              const float* msfty, double rdx, double rdy, const float* rdn, const float* rdnw,
                              double kvdif, const float* xkmhd, pywrf_stream_t s);
/* rk_update_scalar, wrf_dynamics.py:4055-4119 */
int pywrf_rk_update_scalar(const pywrf_grid_t* g, int scs, int sce, float* scalar_1, float* scalar_2,
                           const float* sc_tend, const float* advect_tend, const float* msfty, const float* mu_old,
                           const float* mu_new, const float* mu_base, int rk_step, double dt, int spec_zone,
                           pywrf_stream_t s);
/* Two rk_update_scalar calls that share everything but the two species
 * arrays, their range and their tendencies (the moist and the tracer update
 * of solver.py:1548-1636), in one launch.  Same results as the two calls.  No
 * counterpart in the reference: a batching. */
int pywrf_rk_update_scalar_pair(const pywrf_grid_t* g, int rk_step, int scs, int sce, float* scalar_1,
                                float* scalar_2, const float* sc_tend, const float* advect_tend, int scs2, int sce2,
                                float* scalar_1b, float* scalar_2b, const float* sc_tend_b,
                                const float* advect_tend_b, const float* msfty, const float* mu_old,
                                const float* mu_new, const float* mu_base, double dt, int spec_zone,
                                pywrf_stream_t s);
/* calc_p_rho_phi, wrf_dynamics.py:4124-4164 (phi is not re-integrated, as in the reference) */
int pywrf_calc_p_rho_phi(const pywrf_grid_t* g, const float* moist, float* al, const float* alb, const float* mu,
                         const float* muts, const float* ph, float* p, const float* pb, const float* t, double p0,
                         double t0, const float* rdnw, pywrf_stream_t s);
/* phy_prep_part2, wrf_dynamics.py:4179-4213 */
int pywrf_phy_prep_part2(const pywrf_grid_t* g, const float* mut, float* rthblten, float* rublten,
                         float* rvblten, float* rqvblten, float* rqcblten, float* rqiblten, pywrf_stream_t s);
/* set_w_surface, wrf_dynamics.py:4218-4324 */
int pywrf_set_w_surface(const pywrf_grid_t* g, float* w, const float* ht, const float* u, const float* v,
                        const float* cf, double rdx, double rdy, const float* msftx, const float* msfty,
                        pywrf_stream_t s);

/* ---- lateral boundaries: wrf_boundary.py:52-587 ------------------------ */
/* Boundary arrays in the reference driver's layout (solver.py:468-703): an
 * x-boundary array (bxs, bxe, btxs, btxe) is a contiguous (5, kz, ny) array,
 * the storage of the driver's (kz, ny, 5) view permute(1, 2, 0); a y-boundary
 * array (bys, bye, btys, btye) a contiguous (5, kz, nx) array, the storage of
 * its (kz, 5, nx) view permute(1, 0, 2).  kz is the level count of the field
 * (g->nz); the strip width is the reference's spec_bdy_width = 5. */
/* relax_bdytend, wrf_boundary.py:52-161 (fcx, gcx: 5 floats each) */
int pywrf_relax_bdytend(const pywrf_grid_t* g, const float* field, float* field_tend, const float* bxs,
                        const float* bxe, const float* bys, const float* bye, const float* btxs, const float* btxe,
                        const float* btys, const float* btye, char variable_in, double dtbc, const float* fcx,
                        const float* gcx, pywrf_stream_t s);
/* relax_bdytend with the boundary time dtbc read from device memory (one float)
 * when the kernel runs, so that a captured HIP graph of the step stays valid
 * as dtbc changes from step to step.  Same arithmetic as pywrf_relax_bdytend.
 * No counterpart in the reference (its dtbc is a Python float, solver.py:713). */
int pywrf_relax_bdytend_dtbc_dev(const pywrf_grid_t* g, const float* field, float* field_tend, const float* bxs,
                                 const float* bxe, const float* bys, const float* bye, const float* btxs,
                                 const float* btxe, const float* btys, const float* btye, char variable_in,
                                 const float* dtbc, const float* fcx, const float* gcx, pywrf_stream_t s);
/* spec_bdytend, wrf_boundary.py:188-225 */
int pywrf_spec_bdytend(const pywrf_grid_t* g, float* field_tend, const float* btxs, const float* btxe,
                       const float* btys, const float* btye, char variable_in, pywrf_stream_t s);

/* n consecutive spec_bdytend calls on distinct tendencies (solver.py:
 * 1220-1254) in one launch: tends[f], variables[f] as in pywrf_spec_bdytend,
 * bdy_tends[4f .. 4f+3] its btxs, btxe, btys, btye, is2d[f] selects the index
 * set g2 (a one-level field) over g3.  No counterpart in the reference: a
 * batching. */
int pywrf_spec_bdytend_group(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n, void* const* tends,
                             void* const* bdy_tends, const char* variables, const int* is2d, pywrf_stream_t s);
/* spec_bdyupdate, wrf_boundary.py:228-284 (kds == kde == 0: 2-D field) */
int pywrf_spec_bdyupdate(const pywrf_grid_t* g, float* field, const float* field_tend, double dt, char variable,
                         pywrf_stream_t s);

/* n consecutive spec_bdyupdate calls on distinct fields (as the acoustic step
 * issues them, solver.py:1366-1406) in one launch: fields[f], tends[f],
 * variables[f] as in pywrf_spec_bdyupdate, is2d[f] selects the 2-D index set
 * g2 (kds = kde = 0) over g3.  No counterpart in the reference: a batching. */
int pywrf_spec_bdyupdate_group(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n, void* const* fields,
                               void* const* tends, const char* variables, const int* is2d, double dt,
                               pywrf_stream_t s);
/* spec_bdyupdate_ph, wrf_boundary.py:287-358 */
int pywrf_spec_bdyupdate_ph(const pywrf_grid_t* g, const float* ph_save, float* field, const float* field_tend,
                            const float* mu_tend, const float* muts, double dt, char variable_in,
                            pywrf_stream_t s);
/* zero_grad_bdy, wrf_boundary.py:361-405 */
int pywrf_zero_grad_bdy(const pywrf_grid_t* g, float* field, char variable_in, pywrf_stream_t s);
/* spec_bdyupdate_ph on `field` and zero_grad_bdy on `zg_field` (another
 * array), as the two calls in that order would leave them, in one launch
 * (the substep's strip updates after advance_w, solver.py:1458-1468).  No
 * counterpart in the reference. */
int pywrf_spec_bdyupdate_ph_zero_grad(const pywrf_grid_t* g, const float* ph_save, float* field,
                                      const float* field_tend, const float* mu_tend, const float* muts, double dt,
                                      char variable_in, float* zg_field, char zg_variable, pywrf_stream_t s);
/* mass_weight, wrf_boundary.py:408-415 (tile = its..ite x jts..jte x kts..kte of g) */
int pywrf_mass_weight(const pywrf_grid_t* g, const float* field, const float* mut, float* rfield,
                      pywrf_stream_t s);
/* relax_bdy_scalar, wrf_boundary.py:418-451 */
int pywrf_relax_bdy_scalar(const pywrf_grid_t* g, float* scalar_tend, const float* scalar, const float* mu,
                           const float* bxs, const float* bxe, const float* bys, const float* bye,
                           const float* btxs, const float* btxe, const float* btys, const float* btye, double dtbc,
                           const float* fcx, const float* gcx, pywrf_stream_t s);
/* relax_bdy_scalar with dtbc in device memory (see pywrf_relax_bdytend_dtbc_dev). */
int pywrf_relax_bdy_scalar_dtbc_dev(const pywrf_grid_t* g, float* scalar_tend, const float* scalar, const float* mu,
                                    const float* bxs, const float* bxe, const float* bys, const float* bye,
                                    const float* btxs, const float* btxe, const float* btys, const float* btye,
                                    const float* dtbc, const float* fcx, const float* gcx, pywrf_stream_t s);
/* spec_bdy_scalar, wrf_boundary.py:454-472 */
int pywrf_spec_bdy_scalar(const pywrf_grid_t* g, float* scalar_tend, const float* btxs, const float* btxe,
                          const float* btys, const float* btye, pywrf_stream_t s);
/* flow_dep_bdy, wrf_boundary.py:475-510 */
int pywrf_flow_dep_bdy(const pywrf_grid_t* g, float* field, const float* u, const float* v, pywrf_stream_t s);

/* flow_dep_bdy on n distinct fields with the same u, v (the moisture species
 * loop, solver.py:1599-1602) in one launch.  No counterpart in the reference:
 * a batching. */
int pywrf_flow_dep_bdy_group(const pywrf_grid_t* g, int n, void* const* fields, const float* u, const float* v,
                             pywrf_stream_t s);
/* spec_bdy_final, wrf_boundary.py:513-587 (msf / mu ignored where the reference does not couple) */
int pywrf_spec_bdy_final(const pywrf_grid_t* g, float* field, const float* mu, const float* msf, const float* bxs,
                         const float* bxe, const float* bys, const float* bye, const float* btxs,
                         const float* btxe, const float* btys, const float* btye, char variable_in, double dtbc,
                         pywrf_stream_t s);

/* n consecutive spec_bdy_final calls on distinct fields (solver.py:1863-1903)
 * in one launch: fields[f], mus[f], msfs[f], variables[f] as in
 * pywrf_spec_bdy_final, bdys[8f .. 8f+7] its bxs, bxe, bys, bye, btxs, btxe,
 * btys, btye; one dtbc.  No counterpart in the reference: a batching. */
int pywrf_spec_bdy_final_group(const pywrf_grid_t* g, int n, void* const* fields, void* const* mus,
                               void* const* msfs, void* const* bdys, const char* variables, double dtbc,
                               pywrf_stream_t s);
/* pywrf_spec_bdy_final_group with dtbc in device memory (see
 * pywrf_relax_bdytend_dtbc_dev). */
int pywrf_spec_bdy_final_group_dtbc_dev(const pywrf_grid_t* g, int n, void* const* fields, void* const* mus,
                                        void* const* msfs, void* const* bdys, const char* variables,
                                        const float* dtbc, pywrf_stream_t s);
/* The end-of-step boundary epilogue (solver.py:1776-1910) in one launch:
 * set_physical_bc3d / set_physical_bc2d on n distinct fields (as
 * pywrf_set_physical_bc_group; n <= 24), then spec_bdy_final on nfinal <= 6
 * of them: final_of[m] is the index in fields of the m-th, with mus[m],
 * msfs[m], the eight boundary arrays bdys[8m .. 8m+7] and final_variables[m]
 * as in pywrf_spec_bdy_final_group.  Each (field, level) is one workgroup that
 * fills the plane's halo and then sets its boundary rows: the same results as
 * the two grouped calls in that order.  No counterpart in the reference: a
 * fusion of its calls. */
int pywrf_boundary_epilogue(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n, void* const* fields,
                            const char* variables, const int* is2d, int nfinal, const int* final_of, void* const* mus,
                            void* const* msfs, void* const* bdys, const char* final_variables, double dtbc,
                            pywrf_stream_t s);
/* pywrf_boundary_epilogue with dtbc in device memory. */
int pywrf_boundary_epilogue_dtbc_dev(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n, void* const* fields,
                                     const char* variables, const int* is2d, int nfinal, const int* final_of,
                                     void* const* mus, void* const* msfs, void* const* bdys,
                                     const char* final_variables, const float* dtbc, pywrf_stream_t s);

/* ---- moist physics: wrf_physics.py:27-2137 ------------------------------ */
/* moist_physics_prep_em, wrf_physics.py:27-91 */
int pywrf_moist_physics_prep_em(const pywrf_grid_t* g, const float* t_new, double t0, float* rho, const float* al,
                                const float* alb, const float* p, float* p8w, double p0, const float* pb,
                                const float* ph, const float* phb, float* th_phy, float* pii, float* pf, float* z,
                                float* z_at_w, float* dz8w, float* h_diabatic, const float* qv, float* qv_diabatic,
                                const float* qc, float* qc_diabatic, const float* fzm, const float* fzp,
                                pywrf_stream_t s);
/* microphysics_driver, wrf_physics.py:94-134, forwarding to wsm6 */
int pywrf_microphysics_driver(const pywrf_grid_t* g, float* th, const float* rho, const float* pi_phy,
                              const float* p, const float* dz8w, double dt, float* qv, float* qc, float* qr,
                              float* qi, float* qs, float* qg, float* rainnc, float* rainncv, float* snownc,
                              float* snowncv, float* sr, float* graupelnc, float* graupelncv, pywrf_stream_t s);
/* wsm6 with nislfv_rain_plm / nislfv_rain_plm6 / slope_*, wrf_physics.py:137-1584, 1988-2103 */
int pywrf_wsm6(const pywrf_grid_t* g, float* th, float* q, float* qc, float* qr, float* qi, float* qs, float* qg,
               const float* den, const float* pii, const float* p, const float* delz, double delt, float* rain,
               float* rainncv, float* snow, float* snowncv, float* sr, float* graupel, float* graupelncv,
               pywrf_stream_t s);
/* moist_physics_finish_em, wrf_physics.py:2106-2137 */
int pywrf_moist_physics_finish_em(const pywrf_grid_t* g, float* t_new, const float* th_phy, float* h_diabatic,
                                  double dt, const float* qv, float* qv_diabatic, const float* qc,
                                  float* qc_diabatic, pywrf_stream_t s);

/* ---- WSM6 column surrogate (a learned replacement for wsm6, see below) --- */
/* wsm6 on a device list of the scheme's tile columns, c = (j - jts) *
 * (ite - 1 - its) + (i - its) for c in cols[0 .. *ncols): the other columns are
 * untouched.  The surrogate's fallback; no counterpart in the reference. */
int pywrf_wsm6_columns(const pywrf_grid_t* g, float* th, float* q, float* qc, float* qr, float* qi, float* qs,
                       float* qg, const float* den, const float* pii, const float* p, const float* delz, double delt,
                       float* rain, float* rainncv, float* snow, float* snowncv, float* sr, float* graupel,
                       float* graupelncv, const int* cols, const int* ncols, pywrf_stream_t s);

/* Dense layers on the matrix cores: C[M x N] = A[M x K] B[N x K]^T with BF16
 * A and B (row-major, K contiguous, K a multiple of 32, lda / ldb multiples
 * of 8, 16-byte aligned) and FP32 accumulation.  epi: 0 FP32 C (+ bias[N] if
 * given), 1 BF16 relu(C + bias), 2 BF16 C * [H > 0] (H BF16 [M x ldh]),
 * 3 FP32 atomic accumulation into C (K split over ksplit workgroups; C must
 * start at zero), 4 FP32 C^T (+ bias) into C [N x ldc], ldc >= M.  No
 * counterpart in the reference (the surrogate is new). */
int pywrf_mlp_gemm(int epi, int M, int N, int K, const void* A, int lda, const void* B, int ldb, void* C, int ldc,
                   const float* bias, const void* H, int ldh, int ksplit, pywrf_stream_t s);
/* W [N x K] (BF16, row-major, ldw) rewritten in the B-operand fragment order
 * of v_mfma_f32_32x32x16_bf16: for row tile nt of 32 rows and k-step ks of 16
 * columns, lane l's 8 values W[32 nt + l % 32][16 ks + 8 (l / 32) + 0..7]
 * at out[((nt (K / 16) + ks) 64 + l) 8]; rows at or beyond N are zero.  out
 * holds ceil(N / 32) 32 K BF16.  K a multiple of 16. */
int pywrf_mlp_frag_pack(int N, int K, const void* W, int ldw, void* out, pywrf_stream_t s);
/* out[c][r] = in[r][c], BF16. */
int pywrf_mlp_transpose(int R, int C, const void* in, int ldi, void* out, int ldo, pywrf_stream_t s);
/* out[r] = sum_{c < C} in[r][c] (BF16 in, FP32 out): bias gradients. */
int pywrf_mlp_rowsum(int R, int C, const void* in, int ld, float* out, pywrf_stream_t s);
/* Weighted squared error of Y (FP32 [B x ldy]) against T over N outputs:
 * *loss += sum w[n] (Y - T)^2 scale, dY (BF16 [B x lddy]) = 2 w[n] (Y - T) scale
 * (0 in columns N .. lddy). */
int pywrf_mlp_mse(int B, int N, const float* Y, int ldy, const float* T, int ldt, const float* w, double scale,
                  void* dY, int lddy, float* loss, pywrf_stream_t s);
/* One bias-corrected Adam step of an FP32 [rows x cols] parameter with
 * gradient g * gscale (step number t >= 1); wb / wtb (optional) receive its
 * BF16 copy and transposed BF16 copy. */
int pywrf_mlp_adam(int rows, int cols, float* w, const float* g, float* m, float* v, double lr, double beta1,
                   double beta2, double eps, double gscale, int t, void* wb, void* wtb, pywrf_stream_t s);

/* The trained network: 369 -> 512 -> 512 -> 287, ReLU.  Inputs per column:
 * T = th * pii, p, rho, qv, qc, qr, qi, qs, qg over 41 levels (feature
 * var * 41 + k, padded to 384), standardised with in_mean / in_rstd; outputs:
 * standardised tendencies of T, qv, qc, qr, qi, qs, qg (feature var * 41 + k,
 * 287 used of 288), de-standardised with out_mean / out_std.  qtol: the most
 * negative mixing ratio accepted (smaller ones are clipped to 0); wtol: the
 * largest accepted relative gain of the column's water.  The training
 * envelope: in_lo / in_hi (384 each) bound every input feature the column
 * may show, out_absmax (288) every tendency the network may return, as seen
 * in the training set; a column outside it is not trusted to the network.
 * envelope scales the bounds (the input range is widened by (envelope - 1)
 * times its width on both sides, the tendency bound multiplied by it).
 * dtmax: the largest temperature change (K) the network may make at a level
 * in one call; columns it would heat or cool more (deep convection) are left
 * to the plain scheme. */
typedef struct pywrf_surrogate {
  const void* w1; /* BF16 [512 x 384] */
  const void* w2; /* BF16 [512 x 512] */
  const void* w3; /* BF16 [288 x 512] (row 287 zero) */
  const float* b1;
  const float* b2;
  const float* b3;
  const float* in_mean;
  const float* in_rstd;
  const float* out_mean;
  const float* out_std;
  const float* in_lo;
  const float* in_hi;
  const float* out_absmax;
  float qtol, wtol, envelope, dtmax;
  /* w1, w2, w3 in the matrix cores' fragment order (pywrf_mlp_frag_pack):
   * what the inline network reads */
  const void* w1f;
  const void* w2f;
  const void* w3f;
} pywrf_surrogate_t;

/* Bytes of workspace pywrf_wsm6_surrogate needs for grid g (its physics tile)
 * (the fallback list: the network keeps everything else on chip); `chunk` is
 * checked (>= 128) and otherwise unused. */
int pywrf_wsm6_surrogate_workspace(const pywrf_grid_t* g, int chunk, long long* bytes);
/* microphysics_driver / wsm6 (wrf_physics.py:94-134, 137-1214) replaced by the
 * surrogate: same fields, tile and outputs as pywrf_wsm6; every column that
 * fails the checks (negative mixing ratio, water gain, non-finite value, an
 * input or a tendency outside the training envelope) is recomputed by the
 * plain scheme, and their count is written to *nfallback
 * (device int).  workspace: device memory of at least the size above.  No counterpart in the
 * reference (its README, README.md:170-173, describes but does not ship the
 * learned coupling). */
int pywrf_wsm6_surrogate(const pywrf_grid_t* g, const pywrf_surrogate_t* net, float* th, const float* rho,
                         const float* pii, const float* p, const float* delz, double dt, float* qv, float* qc,
                         float* qr, float* qi, float* qs, float* qg, float* rain, float* rainncv, float* snow,
                         float* snowncv, float* sr, float* graupel, float* graupelncv, void* workspace,
                         long long workspace_bytes, int* nfallback, pywrf_stream_t s);

/* The surrogate network's standardised outputs for every column of g's
 * physics tile (no checks, nothing applied): Y[f * ldy + c], f < 288, c the
 * column (row-major over the tile).  fused = 1: one launch that keeps the
 * inputs and hidden layers on chip (the form pywrf_wsm6_surrogate runs;
 * ldy a multiple of 4, Y 16-byte aligned); fused = 0: the inputs packed to
 * memory and three GEMM launches, through `work` (2816 bytes per column,
 * columns rounded up to 128).  Both give the same bits.  No counterpart in
 * the reference. */
int pywrf_wsm6_surrogate_outputs(const pywrf_grid_t* g, const pywrf_surrogate_t* net, const float* th,
                                 const float* rho, const float* pii, const float* p, const float* qv, const float* qc,
                                 const float* qr, const float* qi, const float* qs, const float* qg, float* Y, int ldy,
                                 int fused, void* work, long long work_bytes, pywrf_stream_t s);

/* pywrf_wsm6_surrogate without a workspace: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation): the
 * buffers are a static device array of 0.26 MB held on every GPU the library
 * runs on until the process ends.  Limit: at most 65536 tile columns (the
 * 240 x 240 domain has 51529); more return PYWRF_E_RANGE.  It calls no
 * allocator.  Calls that overlap on one GPU race (they share its array). */
int pywrf_wsm6_surrogate_noalloc(const pywrf_grid_t* g, const pywrf_surrogate_t* net, float* th, const float* rho,
                                 const float* pii, const float* p, const float* delz, double dt, float* qv,
                                 float* qc, float* qr, float* qi, float* qs, float* qg, float* rain, float* rainncv,
                                 float* snow, float* snowncv, float* sr, float* graupel, float* graupelncv,
                                 int* nfallback, pywrf_stream_t s);
/* The limits of pywrf_wsm6_surrogate_noalloc: columns per call, bytes of its array. */
int pywrf_wsm6_surrogate_noalloc_limits(int* max_columns, long long* scratch_bytes);

/* The NaN check of the reference's loop (solver.py:1648-1654, 1768-1774:
 * torch.isnan(p).any() after each RK stage and after the physics, then exit)
 * kept on the device: scans the grid's nz*ny*nx field and, if it holds a NaN,
 * sets flag[0] to `code` (> 0) unless an earlier check already set it, then
 * flag[1] to the linear index of one NaN point of the check that set flag[0],
 * and adds the NaN count to flag[2].  flag: 4 device ints, cleared by
 * pywrf_nan_flag_clear (a kernel, so that both can sit in a HIP graph).  The
 * host reads the flag when it chooses.  A replacement of the reference's
 * host-side check, not of a function. */
int pywrf_nan_scan(const pywrf_grid_t* g, const float* field, int code, int* flag, pywrf_stream_t s);
int pywrf_nan_flag_clear(int* flag, pywrf_stream_t s);
/* The boundary clock of a captured step (solver.py:457-458, 713, where the
 * host advances it): step (int64) += 1; it_bdy (int64) = floor(step /
 * bdy_interval) when step mod bdy_interval == 1; dtbc (float) = (step -
 * it_bdy * bdy_interval) * dt formed in double, rounded once.  One-value
 * device arrays, advanced by a kernel so that a graph of steps replays it.
 * Not a function of the reference: its driver keeps these on the host. */
int pywrf_clock_advance(void* step, void* it_bdy, float* dtbc, long long bdy_interval, double dt, pywrf_stream_t s);

/* ---------------------------------------------------------------- adjoints
 * Transposes of linear parts of the step, for gradients through it.  The
 * reference has no hand-written adjoints (it relies on torch autograd, and
 * runs its step under no_grad, solver.py:456, 710, 1748); each entry point
 * names the reference function whose linear map it transposes. */

/* Adjoint of advect_scalar (wrf_dynamics.py:1841-2028) with respect to the
 * field, the mass fluxes held fixed: lam_field (+)= A^T lam_tend, where A maps
 * the field to the tendency pywrf_advect_scalar_fresh writes (zero outside
 * the advection box).  accumulate 0: lam_field is overwritten on the whole
 * array; 1: added to. */
int pywrf_advect_scalar_adjoint(const pywrf_grid_t* g, const float* lam_tend, const float* ru, const float* rv,
                                const float* rom, const float* msftx, const float* fzm, const float* fzp, double rdx,
                                double rdy, const float* rdzw, float* lam_field, int accumulate, pywrf_stream_t s);
/* Adjoint of one species of rk_update_scalar (wrf_dynamics.py:4055-4119):
 * given lam_new, the sensitivity of scalar_2 after the update, adds the
 * sensitivity of scalar_1 to lam_n, and writes that of advect_tend (lam_adv),
 * of sc_tend (lam_sc; may be null) and of scalar_2 before the update
 * (lam_prev: lam_new where the update leaves scalar_2 alone, zero where it
 * overwrites it). */
int pywrf_rk_update_scalar_adjoint(const pywrf_grid_t* g, const float* lam_new, const float* msfty,
                                   const float* mu_old, const float* mu_new, const float* mu_base, double dt,
                                   int spec_zone, float* lam_n, float* lam_adv, float* lam_sc, float* lam_prev,
                                   pywrf_stream_t s);
/* Adjoint of advance_w (wrf_dynamics.py:3641-3796) with respect to w, ph,
 * t_2, t_2ave, u, v, ww, muts and muave, every coefficient (mut, c2a, cqw, alt,
 * a, alpha, gamma, the damping from ph_1 and phb) held at the values given:
 * from the sensitivities of the three outputs (t_2ave, w, ph) writes those of
 * w, ph, t_2, t_2ave, u, v (levels 0..2, through the surface w), ww (through
 * the vertical advection of ph_1 + phb) and of the planes muts, muave on the
 * whole array.  t2ave_out, w_out: the forward
 * call's outputs (muts enters by division); ht, msftx, cf, rdx, rdy: its
 * terrain, map factor, surface extrapolation, grid spacing and fnm / fnp
 * arguments.  Also the sensitivities of the forcing rw_tend and ph_tend, of
 * w_save (the damping's reference), of t_1 (in the theta average) and of
 * ph_1 (in the vertical advection of the geopotential, with the call's ww,
 * and in the damping weights, with its w_save), zero outside the levels and
 * columns advanced.
 * work, work2: 3-D scratch arrays of the grid's size. */
int pywrf_advance_w_adjoint(const pywrf_grid_t* g, const float* lam_t2ave_out, const float* lam_w_out,
                            const float* lam_ph_out, const float* mut, const float* muts, const float* t_1,
                            const float* ph_1, const float* phb, const float* c2a, const float* cqw,
                            const float* alt, const float* a, const float* alpha, const float* gamma, double dts,
                            double t0, double epssm, const float* rdnw, const float* rdn, const float* msfty,
                            const float* t2ave_out, const float* w_out, const float* ht, const float* msftx,
                            const float* cf, double rdx, double rdy, const float* fnm, const float* fnp,
                            float* lam_w, float* lam_ph, float* lam_t_2, float* lam_t2ave, float* lam_u,
                            float* lam_v, float* lam_muts, float* lam_muave, float* lam_ww,
                            float* lam_rw_tend, float* lam_ph_tend, float* lam_w_save, float* lam_t_1,
                            const float* ww, const float* w_save, float* lam_ph_1, float* work, float* work2,
                            pywrf_stream_t s);
/* The adjoint of advance_uv (wrf_dynamics.py:3391-3525) with respect to u,
 * v, ru_tend, rv_tend, p, ph, al, mu and mudf (alt, pb, php, cqu, cqv, muu,
 * muv and the map factors held fixed): lam_u_out / lam_v_out are the
 * sensitivities of the returned u and v; the nine outputs are overwritten
 * (lam_mu, lam_mudf are planes).  cf: cf1, cf2, cf3 as three consecutive
 * floats, as for pywrf_advance_uv. */
int pywrf_advance_uv_adjoint(const pywrf_grid_t* g, const float* lam_u_out, const float* lam_v_out,
                             const float* pb, const float* php, const float* alt, const float* muu,
                             const float* cqu, const float* muv, const float* cqv, const float* msfux,
                             const float* msfuy, const float* msfvx, const float* msfvx_inv, const float* msfvy,
                             double rdx, double rdy, double dts, const float* cf, const float* fnm,
                             const float* fnp, double emdiv, const float* rdnw, int spec_zone, float* lam_u,
                             float* lam_v, float* lam_p, float* lam_ph, float* lam_al, float* lam_mu,
                             float* lam_mudf, float* lam_ru_tend, float* lam_rv_tend, pywrf_stream_t s);
/* The adjoint of advance_mu_t (wrf_dynamics.py:3528-3638) with respect to
 * u, v, ww, mu, muave, muts, mudf, t, t_ave, ft, mu_tend and the large-step
 * u_1, v_1, ww_1, t_1 (mut, muu, muv and the map factors held fixed): lam_ww
 * .. lam_mu are the sensitivities of the returned ww, t, t_ave, muave, muts,
 * mudf and mu; ww_out is the returned ww, u and v the call's; the fifteen
 * outputs are overwritten (the mu-like ones are planes). */
int pywrf_advance_mu_t_adjoint(const pywrf_grid_t* g, const float* lam_ww, const float* lam_t,
                               const float* lam_t_ave, const float* lam_muave, const float* lam_muts,
                               const float* lam_mudf, const float* lam_mu, const float* t_1, const float* msftx,
                               const float* msfty, const float* dnw, const float* fnm, const float* fnp,
                               const float* rdnw, double rdx, double rdy, double dts, double epssm,
                               float* lam_ww_ori, float* lam_u, float* lam_v, float* lam_mu_ori,
                               float* lam_muave_ori, float* lam_muts_ori, float* lam_mudf_ori, float* lam_t_ori,
                               float* lam_t_ave_ori, float* lam_ft, float* lam_mu_tend, const float* ww_out,
                               const float* u, const float* v, const float* muu, const float* muv,
                               const float* msfuy, const float* msfvx_inv, float* lam_u_1, float* lam_v_1,
                               float* lam_ww_1, float* lam_t_1, pywrf_stream_t s);
/* The adjoint of calc_p_rho (wrf_dynamics.py:3269-3313) with respect to al,
 * p, pm1, ph, t_2, mu, mut and t_1 (alt and c2a held fixed): lam_al, lam_p,
 * lam_pm1 are the sensitivities of the returned al, p, pm1 (pm1 being the
 * copy of p at step 0); the outputs are overwritten, lam_mu and lam_mut
 * being planes; lam_t_1 may be null. */
int pywrf_calc_p_rho_adjoint(const pywrf_grid_t* g, const float* lam_al, const float* lam_p, const float* lam_pm1,
                             const float* ph, const float* alt, const float* t_2, const float* t_1, const float* c2a,
                             const float* mu, const float* mut, double t0, const float* rdnw, double smdiv, int step,
                             float* lam_al_ori, float* lam_p_ori, float* lam_pm1_in, float* lam_ph, float* lam_t_2,
                             float* lam_mu, float* lam_mut, float* lam_t_1, pywrf_stream_t s);
/* The adjoint of sumflux (wrf_dynamics.py:3799-3839), linear in its fluxes,
 * accumulators and u_lin, v_lin, ww_lin: lam_ru_m .. lam_ww_m are the
 * sensitivities of the accumulators after the call; the outputs (those of
 * ru, rv, ww, of the accumulators before the call and, when lam_u_lin is not
 * null, of u_lin, v_lin, ww_lin, which take muu, muv, msfuy, msfvx_inv) are
 * overwritten. */
int pywrf_sumflux_adjoint(const pywrf_grid_t* g, const float* lam_ru_m, const float* lam_rv_m,
                          const float* lam_ww_m, int iteration, int number_of_small_timesteps, float* lam_ru,
                          float* lam_rv, float* lam_ww, float* lam_ru_m_in, float* lam_rv_m_in, float* lam_ww_m_in,
                          const float* muu, const float* muv, const float* msfuy, const float* msfvx_inv,
                          float* lam_u_lin, float* lam_v_lin, float* lam_ww_lin, pywrf_stream_t s);
/* The adjoint of a run of open-boundary fills (set_physical_bc3d / bc2d,
 * wrf_dynamics.py:164-278), the same arguments as
 * pywrf_set_physical_bc_group with each field replaced by its sensitivity
 * after the fill, which is turned in place into the sensitivity before it. */
int pywrf_set_physical_bc_group_adjoint(const pywrf_grid_t* g3, const pywrf_grid_t* g2, int n,
                                        void* const* lam_fields, const char* variables, const int* is2d,
                                        pywrf_stream_t s);
/* The adjoint of spec_bdyupdate_ph (wrf_boundary.py:287-358) with respect to
 * the field, muts, field_tend and mu_tend (ph_save held fixed): field_in is
 * the field before the update; lam_field holds the sensitivity of the
 * updated field on entry and that of field_in on return; lam_muts and
 * lam_mu_tend are written on the strips' columns only, lam_field_tend and
 * lam_ph_save (which may be null) on the strips' points only (the caller
 * clears them). */
int pywrf_spec_bdyupdate_ph_adjoint(const pywrf_grid_t* g, const float* ph_save, const float* field_in,
                                    const float* field_tend, const float* mu_tend, const float* muts, double dt,
                                    char variable_in, float* lam_field, float* lam_muts, float* lam_field_tend,
                                    float* lam_mu_tend, float* lam_ph_save, pywrf_stream_t s);
/* The adjoint of zero_grad_bdy (wrf_boundary.py:361-405): lam, the
 * sensitivity of the field after the copy, gives lam_in (another array), that
 * of the field before it. */
int pywrf_zero_grad_bdy_adjoint(const pywrf_grid_t* g, const float* lam, char variable_in, float* lam_in,
                                pywrf_stream_t s);
/* The adjoint of small_step_prep (wrf_dynamics.py:3152-3266) past the first
 * RK stage (rk_step > 1; PYWRF_E_ARG otherwise), with pb, p and alt (so c2a)
 * held fixed: from the sensitivities of the saved u..ph and ww, of the
 * acoustic u, v, w, t, ph and of the planes mu_2, muts, muus, muvs, mu_save
 * it writes those of u_2 .. ph_2, u_1 .. ph_1 and ww, and of the planes
 * mu_1 and mu_2.  u_1 .. t_1 and muus, muvs, muts: the call's inputs and
 * results.  work2d: three planes of scratch. */
int pywrf_small_step_prep_adjoint(
    const pywrf_grid_t* g, int rk_step, const float* lam_u_save, const float* lam_v_save, const float* lam_w_save,
    const float* lam_t_save, const float* lam_ph_save, const float* lam_ww_save, const float* lam_u,
    const float* lam_v, const float* lam_w, const float* lam_t, const float* lam_ph, const float* lam_mu_2,
    const float* lam_muts, const float* lam_muus, const float* lam_muvs, const float* lam_mu_save,
    const float* u_1, const float* v_1, const float* w_1, const float* t_1, const float* muus, const float* muvs,
    const float* muts, const float* muu, const float* muv, const float* mut, const float* msfuy,
    const float* msfvx_inv, const float* msfty, float* lam_u_2_in, float* lam_v_2_in, float* lam_w_2_in,
    float* lam_t_2_in, float* lam_ph_2_in, float* lam_u_1, float* lam_v_1, float* lam_w_1, float* lam_t_1,
    float* lam_ph_1, float* lam_ww, float* lam_mu_1, float* lam_mu_2_in, float* work2d, pywrf_stream_t s);
/* The adjoint of calc_mu_uv_1 (wrf_dynamics.py:3844-3860): from the
 * sensitivities of the returned muu, muv, those of mu and of the incoming
 * muu, muv (planes, overwritten). */
int pywrf_calc_mu_uv_1_adjoint(const pywrf_grid_t* g, const float* lam_muu, const float* lam_muv, float* lam_mu,
                               float* lam_muu_in, float* lam_muv_in, pywrf_stream_t s);
/* The adjoint of small_step_finish (wrf_dynamics.py:3863-3978) with mut,
 * muu, muv, the map factors and h_diabatic held fixed: from the
 * sensitivities of the returned u, v, w, ph, ww, t and mu_2 (u_out .. t_out
 * the forward's results) those of the acoustic u_2, v_2, w_2, ph_2, ww, t_2,
 * mu_2, of the saved fields and of the planes muus, muvs, muts, mu_save. */
int pywrf_small_step_finish_adjoint(
    const pywrf_grid_t* g, const float* lam_u, const float* lam_v, const float* lam_w, const float* lam_ph,
    const float* lam_ww, const float* lam_t, const float* lam_mu_2, const float* u_out, const float* v_out,
    const float* w_out, const float* t_out, const float* muu, const float* muus, const float* muv,
    const float* muvs, const float* mut, const float* muts, const float* msfuy, const float* msfvx,
    const float* msfty, float* lam_u_2, float* lam_v_2, float* lam_w_2, float* lam_ph_2, float* lam_ww_in,
    float* lam_t_2, float* lam_u_save, float* lam_v_save, float* lam_w_save, float* lam_ph_save, float* lam_ww1,
    float* lam_t_save, float* lam_muus, float* lam_muvs, float* lam_muts, float* lam_mu_2_in, float* lam_mu_save,
    pywrf_stream_t s);
/* The adjoint of advect_u (wrf_dynamics.py:1112-1307) with respect to u,
 * the mass fluxes held fixed (the edge x-faces' upwind choice by the sign of
 * u taken as in the forward with this u): lam_tendency, the sensitivity of
 * the tendency advect_u adds, gives lam_u (overwritten). */
int pywrf_advect_u_adjoint(const pywrf_grid_t* g, const float* lam_tendency, const float* u, const float* ru,
                           const float* rv, const float* rom, const float* msfux, const float* fzm, const float* fzp,
                           double rdx, double rdy, const float* rdzw, float* lam_u, pywrf_stream_t s);
/* The adjoint of advect_v (wrf_dynamics.py:1332-1532) with respect to v, as
 * pywrf_advect_u_adjoint (the edge y-faces' upwind choice by the sign of v). */
int pywrf_advect_v_adjoint(const pywrf_grid_t* g, const float* lam_tendency, const float* v, const float* ru,
                           const float* rv, const float* rom, const float* msfvx, const float* msfvy,
                           const float* fzm, const float* fzp, double rdx, double rdy, const float* rdzw,
                           float* lam_v, pywrf_stream_t s);
/* The adjoint of advect_w (wrf_dynamics.py:1535-1838) with respect to w, the
 * mass fluxes held fixed: lam_tendency gives lam_w (overwritten). */
int pywrf_advect_w_adjoint(const pywrf_grid_t* g, const float* lam_tendency, const float* ru, const float* rv,
                           const float* rom, const float* msftx, const float* fzm, const float* fzp, double rdx,
                           double rdy, const float* rdzu, float* lam_w, pywrf_stream_t s);
/* The column tridiagonal solve of advance_w (wrf_dynamics.py:3758-3765), in
 * place on w over advance_w's columns, levels 1..kte-1 (level 0 is the lower
 * boundary value): the same operations in the same order. */
int pywrf_tridiag_solve(const pywrf_grid_t* g, const float* a, const float* alpha, const float* gamma, float* w,
                        pywrf_stream_t s);
/* Its adjoint: r is the right-hand side the solve was given, x its solution,
 * lam the sensitivity of x, replaced in place by that of r; grad_a,
 * grad_alpha, grad_gamma receive the sensitivities of the coefficients on
 * the solved columns; work: a 3-D scratch array of the grid's size. */
int pywrf_tridiag_solve_adjoint(const pywrf_grid_t* g, const float* a, const float* alpha, const float* gamma,
                                const float* r, const float* x, float* lam, float* grad_a, float* grad_alpha,
                                float* grad_gamma, float* work, pywrf_stream_t s);

#ifdef __cplusplus
}
#endif
#endif
