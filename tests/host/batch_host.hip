// CPU emulation of the batched launches (htx_batch.hip): the library's own
// layout and table builders (htx_batch_tables.h) over exact-size host
// buffers; the persistent TENO5 launch walked workgroup by workgroup with the
// same tile chunking, each tile's phases run for every thread id in order;
// the batched Hermite launches walked through the same workgroup -> instance
// maps.  Test infrastructure: it checks the batch bookkeeping on the CPU (and,
// built with AddressSanitizer, that no access leaves its buffer).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../hermite-teno_amd/csrc/htx_batch_tables.h"

using namespace htx;

namespace {

void host_dt(Ctl* c, double dx, double dy) {
    double t = c->t, tf = c->tfinal;
    if (t < tf) {
        double sx, sy;
        std::memcpy(&sx, &c->smax_x, 8);
        std::memcpy(&sy, &c->smax_y, 8);
        double dt = c->cfl * std::fmin(dx, dy) / std::fmax(sx, sy);
        dt = t + dt > tf ? tf - t : dt;
        c->dt = dt;
        c->active = 1;
        c->smax_x = c->smax_y = 0ull;
    } else {
        c->dt = 0.0;
        c->active = 0;
    }
}

void host_tadvance(Ctl* c) {
    if (c->active) {
        c->t += c->dt;
        c->nsteps += 1;
    }
}

void atomic_max_bits(unsigned long long* p, double v) {
    unsigned long long b;
    std::memcpy(&b, &v, 8);
    if (b > *p) *p = b;
}

template <int TX, int TY, int FLUX, bool MIXED>
void teno_tile(const Fv2dArgs& a, Fv2dTile<TX, TY>& s, int tile, double dt) {
    using T = Fv2dTile<TX, TY>;
    const int i0 = (tile % a.tiles_x) * TX, j0 = (tile / a.tiles_x) * TY;
    std::vector<double> dxp((size_t)T::NT * T::KC * 4);
    auto dx = [&](int t) { return (double(*)[4]) & dxp[(size_t)t * T::KC * 4]; };
    for (int t = 0; t < T::NT; ++t) ph_load<TX, TY>(a, s, i0, j0, t);
    for (int t = 0; t < T::NT; ++t) ph_recon<TX, TY, false, MIXED>(a, s, t);
    for (int t = 0; t < T::NT; ++t) ph_flux<TX, TY, false, FLUX, MIXED>(a, s, t);
    for (int t = 0; t < T::NT; ++t) ph_xdiv<TX, TY>(a, s, t, dx(t));
    for (int t = 0; t < T::NT; ++t) ph_recon<TX, TY, true, MIXED>(a, s, t);
    for (int t = 0; t < T::NT; ++t) ph_flux<TX, TY, true, FLUX, MIXED>(a, s, t);
    double mx = 0.0, my = 0.0;
    for (int t = 0; t < T::NT; ++t) {
        double smx = 0.0, smy = 0.0;
        ph_epilogue<TX, TY>(a, s, i0, j0, t, (const double(*)[4])dx(t), dt, smx, smy);
        mx = std::fmax(mx, smx);
        my = std::fmax(my, smy);
    }
    if (a.mode == 3) {
        atomic_max_bits(&a.ctl->smax_x, mx);
        atomic_max_bits(&a.ctl->smax_y, my);
    }
}

template <int TX, int TY, int FLUX, bool MIXED>
void teno_launch(const TenoTables& TT, int ninst, int mode, int nblk) {
    using T = Fv2dTile<TX, TY>;
    T* s = new T();
    const int per = (TT.ntiles + nblk - 1) / nblk;
    for (int blk = 0; blk < nblk; ++blk) {
        const int g0 = blk * per, g1 = std::min(g0 + per, TT.ntiles);
        for (int g = g0; g < g1; ++g) {
            const int b = TT.tile_inst[g];
            const Fv2dArgs& a = TT.args[(size_t)(mode - 1) * ninst + b];
            if (!a.ctl->active) continue;
            teno_tile<TX, TY, FLUX, MIXED>(a, *s, g - TT.tile0[b], a.ctl->dt);
        }
    }
    delete s;
}

template <int FLUX>
void herm_launch(const HermTables& H, int ninst, int mode) {
    auto args = [&](int kind, int blk, int& lb) -> const FvhArgs& {
        int b = H.maps[kind][blk];
        lb = blk - H.maps[6 + kind][b];
        return H.args[(size_t)(mode - 1) * ninst + b];
    };
    int lb;
    for (int blk = 0; blk < (int)H.maps[HK_LIMIT].size(); ++blk) {
        const FvhArgs& a = args(HK_LIMIT, blk, lb);
        if (!a.ctl->active) continue;
        // fvh_limit_block: block lb < g0 takes row chains, the rest column chains
        const int g0 = (4 * a.ny + 63) / 64, dir = lb < g0 ? 0 : 1, grp = dir == 0 ? lb : lb - g0;
        const int nl = dir == 0 ? a.ny : a.nx;
        for (int c = 0; c < 64; ++c) {
            int t = grp * 64 + c;
            if (t < 4 * nl) fvh_limit_line(a, dir, t % nl, t / nl);
        }
    }
    for (int y = 0; y < 2; ++y)
        for (int blk = 0; blk < (int)H.maps[y ? HK_RECON_Y : HK_RECON_X].size(); ++blk) {
            const FvhArgs& a = args(y ? HK_RECON_Y : HK_RECON_X, blk, lb);
            if (!a.ctl->active) continue;
            long long nf = y ? (long long)a.nx * (a.ny + 1) : (long long)(a.nx + 1) * a.ny;
            for (int tid = 0; tid < 256; ++tid) {
                long long t = lb * 256ll + tid;
                if (t >= nf) continue;
                const int i = y ? (int)(t % a.nx) : (int)(t % (a.nx + 1)) - 1;
                const int j = y ? (int)(t / a.nx) - 1 : (int)(t / (a.nx + 1));
                if (y && a.mixed) fvh_recon_face<true, true>(a, i, j, a.fy, t, nf);
                else if (y) fvh_recon_face<true, false>(a, i, j, a.fy, t, nf);
                else if (a.mixed) fvh_recon_face<false, true>(a, i, j, a.fx, t, nf);
                else fvh_recon_face<false, false>(a, i, j, a.fx, t, nf);
            }
        }
    for (int y = 0; y < 2; ++y)
        for (int blk = 0; blk < (int)H.maps[y ? HK_FLUX_Y : HK_FLUX_X].size(); ++blk) {
            const FvhArgs& a = args(y ? HK_FLUX_Y : HK_FLUX_X, blk, lb);
            if (!a.ctl->active) continue;
            long long nf = y ? (long long)a.nx * (a.ny + 1) : (long long)(a.nx + 1) * a.ny;
            for (int tid = 0; tid < 64; ++tid) {
                long long t = lb * 64ll + tid;
                if (t >= nf) continue;
                double o[3][12];
                const int i = y ? (int)(t % a.nx) : (int)(t % (a.nx + 1)) - 1;
                const int j = y ? (int)(t / a.nx) - 1 : (int)(t / (a.nx + 1));
                if (y && a.mixed) fvh_flux_face<true, FLUX, true>(a, i, j, a.fy, nf, o);
                else if (y) fvh_flux_face<true, FLUX, false>(a, i, j, a.fy, nf, o);
                else if (a.mixed) fvh_flux_face<false, FLUX, true>(a, i, j, a.fx, nf, o);
                else fvh_flux_face<false, FLUX, false>(a, i, j, a.fx, nf, o);
                double* G = y ? a.gy : a.gx;
                for (int c = 0; c < 12; ++c) G[c * nf + t] = (o[0][c] + o[1][c]) + o[2][c];
            }
        }
    for (int blk = 0; blk < (int)H.maps[HK_UPDATE].size(); ++blk) {
        const FvhArgs& a = args(HK_UPDATE, blk, lb);
        if (!a.ctl->active) continue;
        long long n = (long long)a.nx * a.ny;
        double mx = 0.0, my = 0.0;
        for (int tid = 0; tid < 256; ++tid) {
            long long t = lb * 256ll + tid;
            double smx = 0.0, smy = 0.0;
            if (t < n) fvh_update_cell(a, (int)(t % a.nx), (int)(t / a.nx), a.ctl->dt, smx, smy);
            mx = std::fmax(mx, smx);
            my = std::fmax(my, smy);
        }
        if (a.mode == 3) {
            atomic_max_bits(&a.ctl->smax_x, mx);
            atomic_max_bits(&a.ctl->smax_y, my);
        }
    }
}

template <int TX, int TY>
void teno_dispatch(const TenoTables& TT, int ninst, int mode, int nblk, int flux, int prec) {
    if (flux == FLUX_RUSANOV_CHAR) {
        if (prec == 0) teno_launch<TX, TY, FLUX_RUSANOV_CHAR, false>(TT, ninst, mode, nblk);
        else teno_launch<TX, TY, FLUX_RUSANOV_CHAR, true>(TT, ninst, mode, nblk);
    } else {
        if (prec == 0) teno_launch<TX, TY, FLUX_HLLC, false>(TT, ninst, mode, nblk);
        else teno_launch<TX, TY, FLUX_HLLC, true>(TT, ninst, mode, nblk);
    }
}

// max wave speeds of an instance's state (k_max_speed)
void host_max_speed(const double* q, const htx_params& p, Ctl* c) {
    long long n = (long long)p.nx * p.ny;
    double g1 = p.gamma - 1.0, mx = 0.0, my = 0.0;
    for (long long id = 0; id < n; ++id) {
        double q0 = q[id], q1 = q[id + n], q2 = q[id + 2 * n], q3 = q[id + 3 * n];
        double wu = q1 / q0, wv = q2 / q0, pr = (q3 - 0.5 * (q1 * wu + q2 * wv)) * g1;
        if (p.ref_quirks) {
            mx = std::fmax(mx, max_speed<true>(p.gamma, wu, pr, q0));
            my = std::fmax(my, max_speed<true>(p.gamma, wv, pr, q0));
        } else {
            mx = std::fmax(mx, max_speed<false>(p.gamma, wu, pr, q0));
            my = std::fmax(my, max_speed<false>(p.gamma, wv, pr, q0));
        }
    }
    c->smax_x = c->smax_y = 0ull;
    atomic_max_bits(&c->smax_x, mx);
    atomic_max_bits(&c->smax_y, my);
}

}  // namespace

// nsteps steps of a batch of n instances (TENO5: fp64_fast or fp32_mixed;
// Hermite: fp64_fast),
// tile shape tx x ty of the TENO5 launch, nblk persistent workgroups.  q holds
// every instance's initial state back to back (4 or 12 fields, interior
// cells); on return, the final states, and in t (if not null) each
// instance's time.
extern "C" int batchh_steps(int n, const htx_params* ps, int tx, int ty, int nblk, int nsteps, double* q,
                            double* t) {
    BatchLayout L = make_layout(ps, n);
    int teno_flux = FLUX_HLLC, prec = 0, herm_flux = FLUX_RUSANOV_CHAR;
    for (int i = 0; i < n; ++i) {
        if (ps[i].scheme == HTX_SCHEME_TENO5) {
            teno_flux = ps[i].flux == HTX_FLUX_RUSANOV_CHAR ? FLUX_RUSANOV_CHAR : FLUX_HLLC;
            prec = ps[i].precision == HTX_PREC_FP64_FAST ? 0 : 1;
        } else {
            herm_flux = ps[i].flux == HTX_FLUX_RUSANOV_CHAR ? FLUX_RUSANOV_CHAR
                                                              : (ps[i].ref_quirks ? FLUX_HLLC : FLUX_HLLC_TORO);
        }
    }
    std::vector<double> state(3 * (size_t)L.total, 0.0);
    std::vector<double> scratch((size_t)std::max(herm_scratch_size(L), 1ll), 0.0);
    std::vector<Ctl> ctl(n);
    size_t pos = 0;
    for (int i = 0; i < n; ++i) {
        const htx_params& p = ps[i];
        size_t cnt = (size_t)(p.scheme == HTX_SCHEME_HTENO ? 12 : 4) * p.nx * p.ny;
        std::memcpy(state.data() + L.off[i], q + pos, cnt * sizeof(double));
        pos += cnt;
        ctl[i] = Ctl{};
        ctl[i].t = p.tinit;
        ctl[i].tfinal = p.tmax;
        ctl[i].cfl = p.cfl;
        host_max_speed(state.data() + L.off[i], p, &ctl[i]);
    }
    TenoTables TT = make_teno_tables(L, state.data(), ctl.data(), tx, ty, prec);
    HermTables H = make_herm_tables(L, state.data(), ctl.data(), scratch.data());
    const int nt = (int)L.teno.size(), nh = (int)L.herm.size();
    const int nb = std::max(1, std::min(nblk, TT.ntiles));
    for (int step = 0; step < nsteps; ++step) {
        for (int i = 0; i < n; ++i) host_dt(&ctl[i], L.dxdy[2 * i], L.dxdy[2 * i + 1]);
        for (int mode = 1; mode <= 3; ++mode) {
            if (nt) {
                if (tx == 16 && ty == 16) teno_dispatch<16, 16>(TT, nt, mode, nb, teno_flux, prec);
                else if (tx == 32 && ty == 8) teno_dispatch<32, 8>(TT, nt, mode, nb, teno_flux, prec);
                else if (tx == 8 && ty == 32) teno_dispatch<8, 32>(TT, nt, mode, nb, teno_flux, prec);
                else return -1;
            }
            if (nh) {
                if (herm_flux == FLUX_RUSANOV_CHAR) herm_launch<FLUX_RUSANOV_CHAR>(H, nh, mode);
                else if (herm_flux == FLUX_HLLC) herm_launch<FLUX_HLLC>(H, nh, mode);
                else herm_launch<FLUX_HLLC_TORO>(H, nh, mode);
            }
        }
        for (int i = 0; i < n; ++i) host_tadvance(&ctl[i]);
    }
    pos = 0;
    for (int i = 0; i < n; ++i) {
        const htx_params& p = ps[i];
        size_t cnt = (size_t)(p.scheme == HTX_SCHEME_HTENO ? 12 : 4) * p.nx * p.ny;
        std::memcpy(q + pos, state.data() + L.off[i], cnt * sizeof(double));
        pos += cnt;
        if (t) t[i] = ctl[i].t;
    }
    return 0;
}

#ifdef BATCH_HOST_MAIN
// standalone driver: reads n, the parameter blocks and the states from a
// file, runs the steps, writes the states back
int main(int argc, char** argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s in out tx ty nblk [nsteps]\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int n = 0;
    if (std::fread(&n, sizeof n, 1, f) != 1) return 2;
    std::vector<htx_params> ps(n);
    if (std::fread(ps.data(), sizeof(htx_params), n, f) != (size_t)n) return 2;
    size_t tot = 0;
    for (auto& p : ps) tot += (size_t)(p.scheme == HTX_SCHEME_HTENO ? 12 : 4) * p.nx * p.ny;
    std::vector<double> q(tot);
    if (std::fread(q.data(), sizeof(double), tot, f) != tot) return 2;
    std::fclose(f);
    int rc = batchh_steps(n, ps.data(), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]),
                          argc > 6 ? std::atoi(argv[6]) : 1, q.data(), nullptr);
    f = std::fopen(argv[2], "wb");
    std::fwrite(q.data(), sizeof(double), tot, f);
    std::fclose(f);
    std::printf("batch emulation: %d instances, rc %d\n", n, rc);
    return rc;
}
#endif
