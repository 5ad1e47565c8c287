/*
 * htx_oracle.h -- FP64 CPU restatement of the Hermite-TENO finite-volume
 * residual / time-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libhtx.so) never links or calls it.
 *
 * Every function restates a reference routine; the citation is given at the
 * definition in htx_oracle.c.  State arrays are structure-of-arrays, interior
 * cells only, field-major: q[f*nx*ny + j*nx + i] (x fastest).  Ghost layers
 * are built internally, exactly as the reference's zeroGradBc()/periodicBc().
 */
#ifndef HTX_ORACLE_H
#define HTX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_ZERO_GRAD = 0, ORC_BC_PERIODIC = 1 };
enum { ORC_FLUX_RUSANOV_CHAR = 0, ORC_FLUX_HLLC = 1, ORC_FLUX_CU = 2 /* 1D only */ };

typedef struct {
    int nx, ny;
    double xl, xr, yb, yt;
    double gamma;
    int bc;          /* ORC_BC_* */
    int flux;        /* ORC_FLUX_* */
    int quirks;      /* 1: reproduce reference quirks (HLLC c, HLLC E*, maxSpeed) */
    int recon;       /* 2D characteristic reconstruction: ORC_RECON_* (the reference's is TENO5) */
} orc_params;

enum {
    ORC_RECON_TENO5 = 0, ORC_RECON_TENO5I = 1, ORC_RECON_WENO5JS = 2, ORC_RECON_WENO5Z = 3,
    ORC_RECON_WENO5N = 4, ORC_RECON_TENO5N = 5, ORC_RECON_TENO5A = 6, ORC_RECON_WENO5IS = 7,
    ORC_RECON_WENO5IS_EXT = 8, ORC_RECON_WENO5ISMOD = 9, ORC_RECON_WENO5Z_DIFF = 10, ORC_RECON_TENO5_DIFF = 11,
    ORC_RECON_LINEAR = 12
};
/* the Hermite solver's face reconstruction (orc_fvh2d_*: recon) */
enum { ORC_HRECON_HTENO = 0, ORC_HRECON_HWENO = 1, ORC_HRECON_HTENOI = 2, ORC_HRECON_LINEAR = 3 };

/* ---- scalar building blocks (unit-test surface) ---- */
void orc_teno5(const double u[6], double out[2]);
void orc_linear5(const double u[6], double out[2]);
void orc_teno5_flags(const double u[6], int flags[6]);
void orc_weno5js3q(const double u[5], double out[3]);
void orc_weno5js3q_linear(const double u[5], double out[3]);
void orc_primitive(double gamma, const double q[4], double w[4]);
void orc_conservative(double gamma, const double w[4], double q[4]);
void orc_eigen(double gamma, int dir, const double qa[4], const double qb[4],
               double L[16], double R[16]);
void orc_rusanov_char(double gamma, int dir, const double ql[4], const double qr[4],
                      const double L[16], const double R[16], double f[4]);
void orc_hllc(double gamma, int dir, int quirks, const double ql[4], const double qr[4],
              double f[4]);
double orc_max_speed(double gamma, int dir, int quirks, const double w[4]);
void orc_rk3_combine(int stage, const double* q0, const double* qs, const double* rhs,
                     double dt, double* out, long n);
void orc_gauss9(double nodes[9], double weights[9]);

/* ---- 2D FV (TENO5 + 3-point transverse quadrature) ---- */
void orc_fv2d_rhs(const orc_params* p, const double* q, double* rhs);
/* the same residual with linear (optimal-weight) reconstruction and linear
 * transverse quadrature on every face */
void orc_fv2d_rhs_linear(const orc_params* p, const double* q, double* rhs);
/* TENO5 cut-off decisions per face (label 1 = nothing cut), dir 1 x, 2 y */
void orc_face_labels(const orc_params* p, const double* q, int dir, unsigned char* labels);
void orc_fv2d_max_speed(const orc_params* p, const double* q, double* sx, double* sy);
double orc_fv2d_dt(const orc_params* p, const double* q, double cfl, double t, double tfinal);
void orc_fv2d_rk3(const orc_params* p, double* q, double dt);
int orc_fv2d_advance(const orc_params* p, double* q, double cfl, double t0, double tfinal,
                     int max_steps, double* t_out);

/* ---- 2D Hermite-TENO (FVHermite::System2D): 12 fields u, <dq/dx>, <dq/dy> ---- */
void orc_hteno(double dif, const double u[4], const double d[4], double out[4]);
/* cut-off flags (1 = kept) of hteno's four reconstructions: u^- [0..2],
 * u^+ [3..5], d^- [6..8], d^+ [9..11] */
void orc_hteno_flags(double dif, const double u[4], const double d[4], int flags[12]);
double orc_hermite_limiter(double dif, const double u[3], const double d[3]);
void orc_dflux(double gamma, int dir, const double q[4], const double w[4], double M[16]);
void orc_fvh2d_rhs(const orc_params* p, const double* q, double* rhs);
void orc_fvh2d_rhs_lim(const orc_params* p, const double* q, double* rhs, double* qlim);
void orc_fvh2d_rk3(const orc_params* p, double* q, double dt);
int orc_fvh2d_advance(const orc_params* p, double* q, double cfl, double t0, double tfinal,
                      int max_steps, double* t_out);
void orc_ic_vortex_hermite(const orc_params* p, double cx, double cy, const double* modes,
                           int nmodes, double* q);

/* ---- 1D FV (FV::System1D): 3 fields ---- */
void orc_teno5i(const double u[6], double out[2]);
void orc_weno5(const double u[6], int z, double out[2]); /* z: 0 weno5JS, 1 weno5Z */
/* the remaining Lagrange and Hermite routines (Schemes_Lagrange.cpp:315-1004,
 * :1265-1294; Schemes_Hermite.cpp:234-293, :381-513, :670-767) */
void orc_recon(int recon, const double u[6], double diff, double out[2]);
void orc_e1_cuflux(double gamma, const double L[9], const double R[9], const double ql[3], const double qr[3],
                   double f[3]);
void orc_weno3js(const double u[4], double out[2]);
void orc_weno3js_2q(const double u[3], double out[2]);
void orc_quad2(int kind, const double u[5], double out[2]);          /* 0 JS, 1 Z, 2 TENO */
void orc_quad4_lobatto(int kind, const double u[5], double out[4]);  /* 0 JS, 1 Z, 2 TENO */
void orc_quad4_legendre(int kind, const double u[5], double out[4]); /* 0 JS, 1 Z */
void orc_herm_recon(int kind, double dif, const double u[4], const double d[4], double out[4]);
void orc_hweno_glq(int npts, double dif, const double u[3], const double d[3], double* out); /* npts 2 or 4 */
void orc_e1_hllc(double gamma, int quirks, const double ql[3], const double qr[3], double f[3]);
void orc_fv1d_rhs(int nx, double xl, double xr, double gamma, int recon, int flux, int quirks,
                  const double* q, double* rhs);
int orc_fv1d_advance(int nx, double xl, double xr, double gamma, int recon, int flux, int quirks,
                     double* q, double cfl, double t0, double tfinal, int max_steps, double* t_out);
void orc_ic_sod(int nx, double xl, double xr, double gamma, double* q);
void orc_fv1d_kxrcf(int nx, double xl, double xr, const double* q, unsigned char* flags);
void orc_fv1d_lmpe(int nx, double xl, double xr, int bc, const double* q, unsigned char* flags);

/* ---- test hooks: smoothness indicators ---- */
void orc_teno5_indicators(const double u[6], double SI[6], double b35[2]);
void orc_hteno_indicators(const double u[3], const double ud[3], double SI[3], double SId[3], double brk[2]);

/* ---- initial conditions (restated generators) ---- */
void orc_ic_vortex(const orc_params* p, double cx, double cy, const double* modes,
                   int nmodes, double* q);
void orc_ic_entropy(const orc_params* p, double rho, double u, double v, double pr, double amp, double sx,
                    double sy, int hermite, double* q);

#ifdef __cplusplus
}
#endif
#endif
