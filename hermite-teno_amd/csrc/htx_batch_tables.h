// htx_batch_tables.h -- host-side layout of a batch: the per-instance
// argument blocks of the batched TENO5 and Hermite launches, the global tile
// list of the TENO5 launch and the workgroup -> instance maps of the Hermite
// launches.  Pure host code over caller-supplied base pointers, shared by the
// library (device pointers) and its CPU test harness (host pointers).
#pragma once
#include <vector>

#include "../../include/htx.h"
#include "htx_fv2d_tile.h"
#include "htx_hermite.h"

namespace HTX_NS {

// kinds of batched Hermite launches
constexpr int HK_LIMIT = 0, HK_RECON_X = 1, HK_RECON_Y = 2, HK_FLUX_X = 3, HK_FLUX_Y = 4, HK_UPDATE = 5;
// work items per workgroup of each kind (limiter: 64 chains, row and column
// chains in separate workgroups, lim_blocks)
constexpr int HK_PER_BLOCK[6] = {64, 256, 256, 64, 64, 256};

struct BatchLayout {
    std::vector<htx_params> p;
    std::vector<int> teno, herm;     // instance ids per scheme
    std::vector<long long> off;      // instance -> offset of its state within one state copy
    std::vector<double> dxdy;        // instance -> dx, dy
    long long total = 0;             // doubles per state copy
};

// Layout of the instances' states in a state copy: field-major per
// instance, instance starts aligned to 256 bytes.
inline BatchLayout make_layout(const htx_params* ps, int n) {
    BatchLayout L;
    L.p.assign(ps, ps + n);
    L.off.resize(n);
    L.dxdy.resize(2 * (size_t)n);
    long long total = 0;
    for (int i = 0; i < n; ++i) {
        const htx_params& p = ps[i];
        int nf = p.scheme == HTX_SCHEME_HTENO ? 12 : 4;
        L.off[i] = total;
        total += (long long)nf * p.nx * p.ny;
        total = (total + 31) / 32 * 32;
        L.dxdy[2 * i] = (p.xr - p.xl) / p.nx;
        L.dxdy[2 * i + 1] = (p.yt - p.yb) / p.ny;
        (p.scheme == HTX_SCHEME_TENO5 ? L.teno : L.herm).push_back(i);
    }
    L.total = total;
    return L;
}

struct TenoTables {
    std::vector<Fv2dArgs> args;      // [mode - 1][TENO instance]
    std::vector<int> tile_inst;      // global tile -> TENO instance
    std::vector<int> tile0;          // TENO instance -> first global tile
    int ntiles = 0;
};

// precision codes of a batch's TENO5 instances
enum BatchPrec { BP_FP64_FAST = 0, BP_FP32_MIXED = 1, BP_LEARNED = 2, BP_FP64 = 3, BP_FP32 = 4, BP_LEARNED_FP32 = 5 };

inline int batch_prec_code(int precision) {
    switch (precision) {
        case HTX_PREC_FP64_FAST: return BP_FP64_FAST;
        case HTX_PREC_FP32_MIXED: return BP_FP32_MIXED;
        case HTX_PREC_LEARNED: return BP_LEARNED;
        case HTX_PREC_FP64: return BP_FP64;
        case HTX_PREC_FP32: return BP_FP32;
        case HTX_PREC_LEARNED_FP32: return BP_LEARNED_FP32;
        default: return -1;
    }
}

// Precisions whose TENO5 instances run one stage launch per instance (the
// single-problem kernels: the strict FP64 build, the FP32 bulk build and the
// learned path with its per-step face-flag kernel) instead of the persistent
// launch over all instances' tiles, which is built for fp64_fast and
// fp32_mixed only.
inline bool batch_prec_direct(int code) { return code >= BP_LEARNED; }
inline bool batch_prec_learned(int code) { return code == BP_LEARNED || code == BP_LEARNED_FP32; }

// state = base of the three state copies (A | B | C, each `total` doubles)
// prec: a BatchPrec code; flags_x / flags_y: the learned precisions' face
// flag arrays per TENO5 instance (null otherwise)
inline TenoTables make_teno_tables(const BatchLayout& L, double* state, Ctl* ctl, int TX, int TY, int prec,
                                   unsigned char* const* flags_x = nullptr, unsigned char* const* flags_y = nullptr) {
    TenoTables T;
    const int nt = (int)L.teno.size();
    T.args.resize(3 * (size_t)nt);
    T.tile0.resize(nt);
    QuadConsts qc{};
    fill_quad_consts(qc);
    int g = 0;
    for (int k = 0; k < nt; ++k) {
        const int i = L.teno[k];
        const htx_params& p = L.p[i];
        const int tx = (p.nx + TX - 1) / TX, ty = (p.ny + TY - 1) / TY;
        T.tile0[k] = g;
        for (int t = 0; t < tx * ty; ++t) T.tile_inst.push_back(k);
        g += tx * ty;
        double* A = state + L.off[i];
        double* B = state + L.total + L.off[i];
        double* C = state + 2 * L.total + L.off[i];
        for (int mode = 1; mode <= 3; ++mode) {
            Fv2dArgs a{};
            a.qin = mode == 1 ? A : (mode == 2 ? B : C);
            a.q0 = A;
            a.qout = mode == 1 ? B : (mode == 2 ? C : A);
            a.ctl = ctl + i;
            a.nx = p.nx;
            a.ny = p.ny;
            a.pitch = p.nx;
            a.fstride = (long long)p.nx * p.ny;
            a.bc = p.bc;
            a.mode = mode;
            a.quirks = p.ref_quirks ? 1 : 0;
            a.mixed = (prec != BP_FP64_FAST && prec != BP_FP64) ? 1 : 0;
            a.learned = batch_prec_learned(prec) ? 1 : 0;
            a.bulk32 = (prec == BP_FP32 || prec == BP_LEARNED_FP32) ? 1 : 0;
            a.face_flags_x = flags_x ? flags_x[k] : nullptr;
            a.face_flags_y = flags_y ? flags_y[k] : nullptr;
            a.tiles_x = tx;
            a.tiles_y = ty;
            a.gam = p.gamma;
            a.g1 = p.gamma - 1.0;
            a.dx = L.dxdy[2 * i];
            a.dy = L.dxdy[2 * i + 1];
            a.area = a.dx * a.dy;
            a.qc = qc;
            T.args[(size_t)(mode - 1) * nt + k] = a;
        }
    }
    T.ntiles = g;
    return T;
}

// doubles of Hermite scratch: limited moments (8 fields), face states
// (24 fields per family), face fluxes (12 per family)
inline long long herm_scratch_doubles(const htx_params& p) {
    long long n = (long long)p.nx * p.ny, nxf = (long long)(p.nx + 1) * p.ny, nyf = (long long)p.nx * (p.ny + 1);
    return 8 * n + 36 * (nxf + nyf);
}

inline long long herm_scratch_size(const BatchLayout& L) {
    long long s = 0;
    for (int i : L.herm) s += herm_scratch_doubles(L.p[i]);
    return s;
}

struct HermTables {
    std::vector<FvhArgs> args;   // [mode - 1][Hermite instance]
    std::vector<int> maps[12];   // [kind] workgroup -> instance, [6 + kind] instance -> first workgroup
};

inline HermTables make_herm_tables(const BatchLayout& L, double* state, Ctl* ctl, double* scratch) {
    HermTables H;
    const int nh = (int)L.herm.size();
    H.args.resize(3 * (size_t)nh);
    QuadConsts qc{};
    fill_quad_consts(qc);
    long long so = 0;
    for (int k = 0; k < nh; ++k) {
        const int i = L.herm[k];
        const htx_params& p = L.p[i];
        long long n = (long long)p.nx * p.ny, nxf = (long long)(p.nx + 1) * p.ny, nyf = (long long)p.nx * (p.ny + 1);
        long long items[6] = {4ll * (p.nx + p.ny), nxf, nyf, nxf, nyf, n};
        for (int kind = 0; kind < 6; ++kind) {
            H.maps[6 + kind].push_back((int)H.maps[kind].size());
            long long nb = kind == HK_LIMIT ? lim_blocks(p.nx, p.ny)
                                            : (items[kind] + HK_PER_BLOCK[kind] - 1) / HK_PER_BLOCK[kind];
            for (long long t = 0; t < nb; ++t) H.maps[kind].push_back(k);
        }
        double* S = scratch + so;
        so += herm_scratch_doubles(p);
        FvhArgs a{};
        a.lv = S;
        a.lv_out = S;
        a.fx = S + 8 * n;
        a.fy = a.fx + 24 * nxf;
        a.gx = a.fy + 24 * nyf;
        a.gy = a.gx + 12 * nxf;
        a.ctl = ctl + i;
        a.nx = p.nx;
        a.ny = p.ny;
        a.pitch = p.nx;
        a.fstride = n;
        a.bc = p.bc;
        a.quirks = p.ref_quirks ? 1 : 0;
        a.flux = p.flux;
        a.mixed = (p.precision == HTX_PREC_FP32_MIXED || p.precision == HTX_PREC_FP32) ? 1 : 0;
        a.bulk32 = p.precision == HTX_PREC_FP32 ? 1 : 0;
        a.gam = p.gamma;
        a.g1 = p.gamma - 1.0;
        a.dx = L.dxdy[2 * i];
        a.dy = L.dxdy[2 * i + 1];
        a.area = a.dx * a.dy;
        a.qc = qc;
        fill_herm_consts(a.hc, a.dx);
        double* A = state + L.off[i];
        double* B = state + L.total + L.off[i];
        double* C = state + 2 * L.total + L.off[i];
        for (int mode = 1; mode <= 3; ++mode) {
            FvhArgs m = a;
            m.q = mode == 1 ? A : (mode == 2 ? B : C);
            m.q0 = A;
            m.qout = mode == 1 ? B : (mode == 2 ? C : A);
            m.mode = mode;
            H.args[(size_t)(mode - 1) * nh + k] = m;
        }
    }
    return H;
}

}  // namespace HTX_NS
