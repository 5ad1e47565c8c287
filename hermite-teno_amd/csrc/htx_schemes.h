// htx_schemes.h -- scheme helpers of the host code (htx_scheme, include/htx.h):
// which schemes run the Hermite solver, and the reconstruction codes the
// kernels take (Fv2dArgs / Fv1dArgs::recon: 0 teno5, 1 teno5i, 2 weno5JS,
// 3 weno5Z, RECON_WENO5N .. RECON_LAST of htx_variants.h; FvhArgs::hrecon:
// 0 hteno, 1 hweno, 2 htenoi, 3 linear).
#pragma once
#include "../../include/htx.h"
#include "htx_variants.h"

inline bool htx_is_hermite(int scheme) {
    return scheme == HTX_SCHEME_HTENO || scheme == HTX_SCHEME_HWENO || scheme == HTX_SCHEME_HTENOI ||
           scheme == HTX_SCHEME_HLINEAR;
}

inline int htx_recon_code(int scheme) {
    if (scheme == HTX_SCHEME_TENO5I) return 1;
    if (scheme == HTX_SCHEME_WENO5JS) return 2;
    if (scheme == HTX_SCHEME_WENO5Z) return 3;
    // HTX_SCHEME_WENO5N .. HTX_SCHEME_LINEAR in the order of RECON_WENO5N .. RECON_LINEAR
    static_assert(HTX_SCHEME_LINEAR - HTX_SCHEME_WENO5N == HTX_NS::RECON_LINEAR - HTX_NS::RECON_WENO5N, "codes");
    if (scheme >= HTX_SCHEME_WENO5N && scheme <= HTX_SCHEME_LINEAR)
        return HTX_NS::RECON_WENO5N + (scheme - HTX_SCHEME_WENO5N);
    return 0;
}

inline int htx_hrecon_code(int scheme) {
    return scheme == HTX_SCHEME_HWENO ? 1 : scheme == HTX_SCHEME_HTENOI ? 2 : scheme == HTX_SCHEME_HLINEAR ? 3 : 0;
}
