// htx_strict.hip -- the FP64 drop-in build of the stage kernels: the same
// sources as the fast build (htx_fv2d.hip, htx_fvh2d.hip), compiled with FMA
// contraction off (-ffp-contract=off on this file) and with every division
// of the reference kept as a division.  Every operation is then one IEEE
// operation in the reference's order, so the residual and each RK stage are
// bit-identical to an IEEE evaluation of the reference formulas, and TENO's
// cut-off decisions never diverge from the CPU solver's.
#define HTX_NS htx_strict
#define HTX_EXACT_DIVISION 1
#include "htx_fv2d.hip"
#include "htx_fv1d.hip"

// entry points for htx_api.hip, which only sees the fast build's types (the
// argument structs have the same layout in both builds)
hipError_t htx_strict_launch_fv2d_stage(const void* a, int flux, int tile_x, int tile_y, hipStream_t s) {
    return htx_strict::launch_fv2d_stage(*static_cast<const htx_strict::Fv2dArgs*>(a), flux, tile_x, tile_y, s);
}

// the CFL maxima of a 2D state with the strict build's sound speed (the
// drop-in's first dt after a state is set); ctl points to a Ctl
hipError_t htx_strict_launch_max_speed(const double* q, int nx, int ny, int pitch, long long fstride, double gam,
                                       int quirks, void* ctl, hipStream_t s) {
    return htx_strict::launch_max_speed(q, nx, ny, pitch, fstride, gam, quirks, static_cast<htx_strict::Ctl*>(ctl),
                                        s);
}


// the 1D path (built here only); a points to an Fv1dArgs, whose layout does
// not depend on the build
hipError_t htx_strict_launch_fv1d_stage(const void* a, hipStream_t s) {
    return htx_strict::launch_fv1d_stage(*static_cast<const htx_strict::Fv1dArgs*>(a), s);
}

hipError_t htx_strict_launch_fv1d_max_speed(const double* q, int nx, double gam, void* ctl, hipStream_t s) {
    return htx_strict::launch_fv1d_max_speed(q, nx, gam, static_cast<htx_strict::Ctl*>(ctl), s);
}

hipError_t htx_strict_launch_fv1d_kxrcf(const double* q, int nx, int bc, double radius, unsigned char* flags,
                                        hipStream_t s) {
    return htx_strict::launch_fv1d_kxrcf(q, nx, bc, radius, flags, s);
}

hipError_t htx_strict_launch_fv1d_lmpe(const double* q, int nx, int bc, double dx, unsigned char* flags,
                                       hipStream_t s) {
    return htx_strict::launch_fv1d_lmpe(q, nx, bc, dx, flags, s);
}
