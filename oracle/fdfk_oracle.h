/*
 * CPU oracle for the FDFK2D time-stepping path.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker / CPU baseline.  The product library
 * (fdfk2d_amd/csrc) never links or calls it.
 *
 * Two kinds of functions live here:
 *  (1) literal restatements of source-visible reference routines
 *      (src/Submain.f90: fd_coef0/1/2, check_stability, diff1, diffx1,
 *      diff{x2,z2,zx,xz}_{reg,pml}, loadseis; src/Input.f90 material law;
 *      src/calc_backazimuth.f90 + coord_transform.f90 +
 *      define_rotation_matrix.f90 geometry).  Their FP32/FP64 mixing follows
 *      the Fortran semantics so they agree bit-for-bit with the reference
 *      routines compiled from source (oracle/_ref, built by oracle/Makefile).
 *  (2) this project's formulation of the parts the reference ships only as
 *      closed binaries (updatedispl: PML, free surface, hybrid TF/SF
 *      injection, leapfrog; comput_background_wavefield: FK plane-wave
 *      background).  Those have no reference source to follow; they are
 *      pinned by closed forms and invariants in tests/ (energy decay,
 *      PML reflection, hybrid self-consistency, analytic half-space and
 *      layer-over-half-space responses).
 *
 * Array layouts are the reference's Fortran ones (column-major, z fastest):
 *   c(1:3, izt:nz+npml, 1-npml:nx+npml), ux/uz(1:2, izt:nz+npml, 1-npml:nx+npml)
 *   coef(1:norder, 1:norder)  -> coef[(j-1) + norder*(i-1)]
 */
#ifndef FDFK_ORACLE_H
#define FDFK_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* ---- (1) literal restatements: src/Submain.f90 ---- */
void orc_fd_coef0(int norder, double *coef);               /* Submain.f90:1111-1153 */
void orc_fd_coef1(int norder, double *coef);               /* Submain.f90:1157-1198 */
void orc_fd_coef2(int norder, double *coef);               /* Submain.f90:1202-1243 */
/* returns 0 ok, 1 unstable (reference stops), 2 warning (dt < 0.4 dt_stable);
 * dt_stable written to *dt_stable.  Submain.f90:118-164 */
int orc_check_stability(int nx, int nz, int npml, int izt, int norder, float dx, float dz,
                        double dt, const double *coef2, const float *c, float *dt_stable);
float orc_diff1(int norder, int dire, int nx, int nz, int npml, int izt, int ix, int iz,
                float d, const double *coef, const float *u);           /* :167-212 */
float orc_diffx1(int norder, int nx, int npml, int ix, float dx, const double *coef,
                 const float *u);                                        /* :215-249 */
/* operators on 2-D (izt:nz+npml, 1-npml:nx+npml) arrays; out fully overwritten */
void orc_diffzx_reg(int norder, int nx, int nz, int npml, int izt, float dx, float dz,
                    const double *coef, const float *mat, const float *u, float *out);
void orc_diffzx_pml(int norder, int nx, int nz, int npml, int izt, float dx, float dz,
                    const double *coef, const float *mat, const float *u, float *out);
void orc_diffxz_reg(int norder, int nx, int nz, int npml, int izt, float dx, float dz,
                    const double *coef, const float *mat, const float *u, float *out);
void orc_diffxz_pml(int norder, int nx, int nz, int npml, int izt, float dx, float dz,
                    const double *coef, const float *mat, const float *u, float *out);
void orc_diffx2_reg(int norder, int nx, int nz, int npml, int izt, float dx,
                    const double *coef, const float *mat, const float *u, float *out);
void orc_diffx2_pml(int norder, int nx, int nz, int npml, int izt, float dx,
                    const double *coef, const float *mat, const float *u, float *out);
void orc_diffz2_reg(int norder, int nx, int nz, int npml, int izt, float dz,
                    const double *coef, const float *mat, const float *u, float *out);
void orc_diffz2_pml(int norder, int nx, int nz, int npml, int izt, float dz,
                    const double *coef, const float *mat, const float *u, float *out);
/* Submain.f90:33-116; ux/uz in (1:2,...) layout, level 1 is read */
void orc_loadseis(int nx, int nz, int npml, int izt, float x0, float z0, float dx, float dz,
                  float rx, float rz, const float *ux, const float *uz, float *dsx, float *dsz);
/* Input.f90:364-378: c(2)=vp,c(3)=vs on input -> rho, lambda, mu */
void orc_material_law(long ncell, float *c);

/* geometry: calc_backazimuth.f90:72-197 (backaz), :201-267, define_rotation_matrix.f90,
 * coord_transform.f90:1-61 */
void orc_backaz(double evla, double evlo, double stla, double stlo, double *baz);
void orc_backazimuth_grids(double lat_org, double lon_org, double az_org, int n,
                           const float *coordx, double evla, double evlo, double *bazs);

/* ---- (2) this project's formulation of updatedispl (68 arguments, the
 * reference call at src/FDFK2D.f90:516-522; all by reference) ---- */
void orc_updatedispl(const int *norder, const int *nx, const int *nz, const int *npml,
    const int *izt, const int *is_pml_top, const int *npower, const int *nbd, const int *it,
    const int *nt, const float *dx, const float *dz, const double *dt, const float *vpml,
    const double *R0,
    const float *uxl_reg, const float *uzl_reg, const float *uxr_reg, const float *uzr_reg,
    const float *uxb_reg, const float *uzb_reg, const float *uxl_pml, const float *uzl_pml,
    const float *uxr_pml, const float *uzr_pml, const float *uxb_pml, const float *uzb_pml,
    const float *c, const double *coef0, const double *coef1, float *ux, float *uz,
    float *bpxlx, float *bpzlx, float *bpxrx, float *bpzrx, float *bpxtx, float *bpxtz,
    float *bpztx, float *bpztz, float *bpxbx, float *bpxbz, float *bpzbx, float *bpzbz,
    float *bux1l, float *bux2l, float *bux3l, float *buz1l, float *buz2l, float *buz3l,
    float *bux1r, float *bux2r, float *bux3r, float *buz1r, float *buz2r, float *buz3r,
    float *bux1b, float *bux2b, float *bux3b, float *buz1b, float *buz2b, float *buz3b,
    float *bux1t, float *bux2t, float *bux3t, float *buz1t, float *buz2t, float *buz3t);

/* Pointwise evaluation of the six right-hand-side parts at one cell:
 * parts = {Xx, Zx, Mx, Xz, Zz, Mz} (x-terms, z-terms, mixed terms of the ux
 * and uz equations).  ux2d/uz2d/c are 2-D (NZ x NX) arrays; c3 holds
 * rho, lambda, mu as 3 planes.  Exposed for tests. */
void orc_eval_parts(int norder, int nx, int nz, int npml, int izt, float dx, float dz,
                    const double *coef0, const double *coef1, const float *lam2d,
                    const float *mu2d, const float *ux2d, const float *uz2d,
                    int iz, int ix, float parts[6]);

/* damping profile used by the PML formulation (per padded column / row) */
void orc_damping(int nx, int nz, int npml, int nbd, int is_pml_top, int npower, float dx,
                 float dz, float vpml, double R0, float *dpx /* NX */, float *dpz /* NZ */);

/* ---- FK background (this project's formulation; FP64 inside) ----
 * Same 32 arguments as the reference call src/FDFK2D.f90:432-434. */
void orc_comput_background_wavefield(const int *nx, const int *nz, const int *npml,
    const int *izt, const int *nbd, const int *nt, const int *nlayer_left,
    const int *nlayer_right, const int *src_type, const float *x0, const float *z0,
    const float *dx, const float *dz, const double *dt, const float *ds, const float *f0,
    const double *rayp, const double *bazs, const float *fk_left, const float *fk_right,
    float *uxl_reg, float *uzl_reg, float *uxr_reg, float *uzr_reg, float *uxb_reg,
    float *uzb_reg, float *uxl_pml, float *uzl_pml, float *uxr_pml, float *uzr_pml,
    float *uxb_pml, float *uzb_pml);

/* FK response of one 1-D model at depths z[0..nzq) for one column offset:
 * ux/uz (nzq x ntq, time fastest) of the plane wave at x-offset 0 with the
 * incident wave reaching depth zref at time t0 (seconds).  Exposed for tests. */
int orc_fk_traces(int nlayer, const float *fkmodel, int src_is_p, double p, float f0, float ds,
                  double dt, int ntq, int nzq, const double *z, double zref, double t0,
                  double *ux, double *uz);

#ifdef __cplusplus
}
#endif
#endif
