// Drives the host code of libtetemoko_hip -- island, tile, strip and exchange lists of levels,
// boundary pieces, the coarse-fine fill, flux-register, reflux and average-down lists of
// hierarchies and hierarchy batches, Berger-Rigoutsos box generation, regrid lists, the
// drop-in's tile lists and host conversions, stations and tags -- over seeded random layouts,
// built with AddressSanitizer and UndefinedBehaviorSanitizer against the host-memory stand-in
// of the HIP runtime (hip_host_stand_in.cpp).  No kernel runs: the sanitizers check every host
// memory access, and the driver checks that every call returns TMK_OK (or the documented
// error for the layouts it builds to be refused).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "tetemoko_hip.h"

#define CHECK(x)                                                                          \
    do {                                                                                  \
        const int e_ = (x);                                                               \
        if (e_ != TMK_OK) {                                                               \
            std::fprintf(stderr, "%s:%d: %s -> %d (%s)\n", __FILE__, __LINE__, #x, e_,    \
                         tmk_strerror(e_));                                               \
            std::exit(1);                                                                 \
        }                                                                                 \
    } while (0)

static std::mt19937 rng(20261203);
static int uni(int lo, int hi) { return std::uniform_int_distribution<int>(lo, hi)(rng); }

static tmk_params params(int nx, int ny, int fault, int plastic, int pred, double dx)
{
    tmk_params p{};
    const double cs = 3464.0, cp = 6000.0;
    p.cs = cs; p.cp = cp; p.G = 2.67e-3 * cs * cs;
    p.sxx0 = p.syy0 = p.szz0 = -120.0; p.sxy0 = 70.0;
    p.use_plasticity = plastic;
    p.p_mu = 0.6; p.p_beta = 0.3; p.p_eta = 0.03 * p.G; p.p_b = 1.0;
    p.bc[0] = 0; p.bc[1] = uni(0, 1); p.bc[2] = 0; p.bc[3] = uni(0, 1);
    p.fault_kind = fault;
    p.r0 = 1500.0; p.nsig = 25.0; p.rs_a = 0.014; p.rs_b = 0.018; p.rs_V0 = 1e-6; p.rs_f0 = 0.6; p.rs_L = 0.02;
    p.rs_fw = 0.13; p.rs_Vw = 0.17; p.rs_fexp = -1.0; p.x0 = nx * dx / 2; p.rs_rupture_thr = 1e-3;
    p.ramp_w = 1e40; p.rs_psi0 = 0.75; p.state_law = 1;
    p.fs = 0.677; p.fd = 0.525; p.dw = 0.4; p.sw_rupture_thr = 1e-3;
    p.num_patches = 1; p.xc[0] = nx * dx / 2; p.xw[0] = 1500.0; p.tau_p[0] = 81.6;
    p.fric_width[0] = 1e40; p.fric_width[1] = 1e40; p.outside_f = 10.0;
    p.normal_pred = pred; p.fourth_order = 1; p.prim_limiting = 1; p.char_limiting = 0;
    p.dx = dx;
    p.domain.lo[0] = 0; p.domain.lo[1] = 0; p.domain.hi[0] = nx - 1; p.domain.hi[1] = ny - 1;
    return p;
}

static tmk_box box(int x0, int y0, int x1, int y1)
{
    tmk_box b;
    b.lo[0] = x0; b.lo[1] = y0; b.hi[0] = x1; b.hi[1] = y1;
    return b;
}

// the domain cut at random x and y positions (boxes from 1 cell up), or a random subset of them
static std::vector<tmk_box> partition(int nx, int ny, int ncut, bool subset)
{
    std::vector<int> xs{0, nx}, ys{0, ny};
    for (int k = 0; k < ncut; ++k) { xs.push_back(uni(1, nx - 1)); ys.push_back(uni(1, ny - 1)); }
    std::sort(xs.begin(), xs.end());
    std::sort(ys.begin(), ys.end());
    xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
    ys.erase(std::unique(ys.begin(), ys.end()), ys.end());
    std::vector<tmk_box> out;
    for (size_t j = 0; j + 1 < ys.size(); ++j)
        for (size_t i = 0; i + 1 < xs.size(); ++i)
            if (!subset || uni(0, 2) > 0) out.push_back(box(xs[i], ys[j], xs[i + 1] - 1, ys[j + 1] - 1));
    if (out.empty()) out.push_back(box(0, 0, nx - 1, ny - 1));
    return out;
}

static size_t ghosted(const tmk_box& b)
{
    return (size_t)(b.hi[0] - b.lo[0] + 1 + 2 * TMK_GHOST) * (b.hi[1] - b.lo[1] + 1 + 2 * TMK_GHOST);
}

static void fill_state(std::vector<double>& U, size_t plane, const tmk_params& p)
{
    const double bg[11] = {0, 0, 0, p.sxx0, p.syy0, p.szz0, p.sxy0, p.sxz0, p.syz0, 0, 0};
    U.assign(plane * TMK_NCOMP, 0.0);
    for (int s = 0; s < 2; ++s)
        for (int c = 0; c < 11; ++c)
            for (size_t k = 0; k < plane; ++k)  // mode II: the anti-plane planes stay zero
                U[k + (size_t)(11 * s + c) * plane] = bg[c] + (c == 2 || c == 7 || c == 8 ? 0.0 : 1e-3 * (double)(k % 7));
}

// one level: every update shape, steps, uploads and downloads, stations, tags
static void level_case(int precision)
{
    const int nx = uni(24, 160), ny = uni(16, 96);
    const int fault = uni(0, 1), plastic = uni(0, 1), pred = uni(0, 2);
    const tmk_params p = params(nx, ny, fault, plastic, pred, 25.0);
    tmk_ctx* c = nullptr;
    CHECK(tmk_ctx_create(&c, &p, precision));
    const std::vector<tmk_box> boxes = partition(nx, ny, uni(0, 6), uni(0, 3) == 0);
    tmk_level* L = nullptr;
    CHECK(tmk_level_create(c, (int)boxes.size(), boxes.data(), &L));
    const int nb = fault == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
    for (size_t k = 0; k < boxes.size(); ++k) {
        std::vector<double> U;
        fill_state(U, ghosted(boxes[k]), p);
        std::vector<double> B((size_t)(boxes[k].hi[0] - boxes[k].lo[0] + 1 + 2 * TMK_GHOST) * nb, 0.0);
        CHECK(tmk_level_upload(L, (int)k, U.data(), boxes[k].lo[1] == 0 ? B.data() : nullptr));
    }
    const int shapes[][3] = {{4, 0, 0}, {8, 0, 0}, {32, 0, 0}, {64, 0, 0}, {0, 16, 16}, {0, 32, 8},
                             {0, 32, 12}, {0, 32, 16}, {0, 32, 32}, {0, 64, 4}, {0, 64, 8}, {0, 64, 16}};
    double t = 0.0, dt = 1e-3, dtn = 0.0;
    for (const auto& s : shapes) {
        CHECK(tmk_level_set_update_shape(L, s[0], s[1], s[2]));
        CHECK(tmk_level_step(L, t, dt, &dtn));
        CHECK(tmk_level_run(L, 3, t, dt));
        CHECK(tmk_level_sync(L));
    }
    tmk_profile prof;
    CHECK(tmk_level_profile_kernels(L, 2, t, dt, &prof));
    tmk_tuning tun;
    CHECK(tmk_level_autotune(L, t, dt, 2, &tun));
    for (size_t k = 0; k < boxes.size(); ++k) {
        std::vector<double> U(ghosted(boxes[k]) * TMK_NCOMP);
        std::vector<double> B((size_t)(boxes[k].hi[0] - boxes[k].lo[0] + 1 + 2 * TMK_GHOST) * nb);
        CHECK(tmk_level_download(L, (int)k, U.data(), boxes[k].lo[1] == 0 ? B.data() : nullptr));
    }
    std::vector<unsigned char> tags((size_t)nx * ny);
    tmk_tagparams tp{};
    tp.refine_thresh = 0.01; tp.plastic_thresh = 1e-5; tp.tag_buffer = uni(0, 3);
    tp.x_coarsen[0] = tp.y_coarsen[0] = tp.x_coarsen2[0] = tp.y_coarsen2[0] = -1e40;
    tp.x_coarsen[1] = tp.y_coarsen[1] = tp.x_coarsen2[1] = tp.y_coarsen2[1] = 1e40;
    tp.width_coarsen = 1e40; tp.coarsen2_max_level = 3; tp.stress_thresh = 0.1; tp.sliprate_thresh = 1e-3;
    CHECK(tmk_level_tag(L, &tp, tags.data()));
    bool fault_box = false;
    for (const tmk_box& b : boxes) fault_box = fault_box || b.lo[1] == 0;
    if (fault_box) {
        std::vector<double> xs;
        for (int k = 0; k < uni(1, 9); ++k) xs.push_back((uni(0, nx - 2) + 0.25) * p.dx);
        tmk_stations* st = nullptr;
        const int e = tmk_stations_create(L, (int)xs.size(), xs.data(), 8, &st);
        if (e == TMK_OK) {
            CHECK(tmk_level_run(L, 5, t, dt));
            std::vector<double> rec((size_t)8 * xs.size() * (1 + nb));
            int n = 0;
            CHECK(tmk_stations_read(st, rec.data(), 8, &n));
            CHECK(tmk_stations_destroy(st));
        } else if (e != TMK_E_ARG) {
            CHECK(e);  // stations only on faces of the level's fault boxes
        }
    }
    CHECK(tmk_level_destroy(L));
    CHECK(tmk_ctx_destroy(c));
}

// the drop-in on boxes anywhere in the domain (fault, corners, interior), both precisions
static void patch_case(int precision)
{
    const int nx = uni(40, 120), ny = uni(30, 80);
    const int fault = uni(0, 1);
    const tmk_params p = params(nx, ny, fault, uni(0, 1), uni(0, 2), 25.0);
    tmk_ctx* c = nullptr;
    CHECK(tmk_ctx_create(&c, &p, precision));
    const int nb = fault == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
    for (int k = 0; k < 6; ++k) {
        const int x0 = uni(0, nx - 2), y0 = uni(0, 1) ? 0 : uni(0, ny - 2);
        const tmk_box v = box(x0, y0, uni(x0, std::min(nx - 1, x0 + 70)), uni(y0, std::min(ny - 1, y0 + 50)));
        const tmk_box ub = box(v.lo[0] - TMK_GHOST, v.lo[1] - TMK_GHOST, v.hi[0] + TMK_GHOST, v.hi[1] + TMK_GHOST);
        std::vector<double> U;
        fill_state(U, ghosted(v), p);
        const int vnx = v.hi[0] - v.lo[0] + 1, vny = v.hi[1] - v.lo[1] + 1;
        std::vector<double> F0((size_t)(vnx + 1) * vny * TMK_NCOMP), F1((size_t)vnx * (vny + 1) * TMK_NCOMP);
        const tmk_box bb = box(v.lo[0] - TMK_GHOST, 0, v.hi[0] + TMK_GHOST, 0);
        std::vector<double> B((size_t)(vnx + 2 * TMK_GHOST) * nb, 0.0);
        double* Bp = v.lo[1] == 0 ? B.data() : nullptr;
        double mws = 0.0;
        CHECK(tmk_patch_update_state(c, v, U.data(), ub, TMK_NCOMP, uni(0, 1) ? F0.data() : nullptr,
                                     uni(0, 1) ? F1.data() : nullptr, Bp, bb, nb, 0.0, 1e-3, &mws));
        const int e = tmk_patch_update_state_noalloc(c, v, U.data(), ub, TMK_NCOMP, F0.data(), F1.data(), Bp, bb, nb,
                                                     0.0, 1e-3, &mws);
        if (e != TMK_E_SCRATCH) CHECK(e);
        CHECK(tmk_patch_plastic_update(c, v, U.data(), ub, TMK_NCOMP, 1e-3));
    }
    CHECK(tmk_ctx_destroy(c));
}

// box lists of the next finer level from random tag blobs on a level of nx x ny cells
static std::vector<tmk_box> finer_boxes(int nx, int ny, const std::vector<tmk_box>& coarse, int r)
{
    std::vector<unsigned char> tags((size_t)nx * ny, 0);
    for (int b = 0; b < uni(1, 4); ++b) {
        const int cx = uni(0, nx - 1), cy = uni(0, ny - 1), rx = uni(1, nx / 4 + 1), ry = uni(1, ny / 4 + 1);
        for (int j = std::max(0, cy - ry); j <= std::min(ny - 1, cy + ry); ++j)
            for (int i = std::max(0, cx - rx); i <= std::min(nx - 1, cx + rx); ++i) tags[(size_t)j * nx + i] = 1;
    }
    tmk_meshparams mp{0.7, 2 * r, 8 * r * uni(1, 4), r, 0};
    std::vector<tmk_box> out(4096);
    int n = 0;
    CHECK(tmk_make_boxes(nx, ny, tags.data(), (int)coarse.size(), coarse.empty() ? nullptr : coarse.data(), &mp,
                         (int)out.size(), out.data(), &n));
    out.resize(n);
    return out;
}

struct Hier {
    std::vector<tmk_ctx*> ctx;
    std::vector<tmk_level*> lev;
    std::vector<std::vector<tmk_box>> boxes;
};

static Hier make_hierarchy(int nx, int ny, int nlev, int r, int precision, int fault)
{
    Hier h;
    double dx = 100.0;
    int lnx = nx, lny = ny;
    for (int l = 0; l < nlev; ++l) {
        std::vector<tmk_box> b = l == 0 ? partition(nx, ny, uni(0, 3), false)
                                        : finer_boxes(lnx / r, lny / r, h.boxes[l - 1], r);
        if (b.empty()) break;
        const tmk_params p = params(lnx, lny, fault, 1, TMK_PRED_PPM, dx);
        tmk_ctx* c = nullptr;
        CHECK(tmk_ctx_create(&c, &p, precision));
        tmk_level* L = nullptr;
        CHECK(tmk_level_create(c, (int)b.size(), b.data(), &L));
        const double bg[22] = {0, 0, 0, p.sxx0, p.syy0, p.szz0, p.sxy0, 0, 0, 0, 0,
                               0, 0, 0, p.sxx0, p.syy0, p.szz0, p.sxy0, 0, 0, 0, 0};
        const int nb = fault == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
        for (size_t k = 0; k < b.size(); ++k) {
            std::vector<double> B((size_t)(b[k].hi[0] - b[k].lo[0] + 1 + 2 * TMK_GHOST) * nb, 0.0);
            CHECK(tmk_level_fill(L, (int)k, bg, b[k].lo[1] == 0 ? B.data() : nullptr));
        }
        h.ctx.push_back(c);
        h.lev.push_back(L);
        h.boxes.push_back(b);
        dx /= r;
        lnx *= r;
        lny *= r;
    }
    return h;
}

static void free_hierarchy(Hier& h)
{
    for (tmk_level* L : h.lev) CHECK(tmk_level_destroy(L));
    for (tmk_ctx* c : h.ctx) CHECK(tmk_ctx_destroy(c));
}

// hierarchies from tag-driven layouts: the multi-level lists, a regrid of the finest level,
// and a hierarchy batch of two members
static void hierarchy_case(int precision)
{
    const int r = uni(0, 1) ? 2 : 4;
    const int nx = 8 * r * uni(2, 6), ny = 8 * r * uni(1, 3);
    const int fault = uni(0, 1);
    Hier h = make_hierarchy(nx, ny, uni(2, 3), r, precision, fault);
    const int nlev = (int)h.lev.size();
    if (nlev < 2) { free_hierarchy(h); return; }
    tmk_amr* A = nullptr;
    CHECK(tmk_amr_create(nlev, h.lev.data(), r, &A));
    CHECK(tmk_amr_run(A, 3, 0.0, 1e-3));
    CHECK(tmk_amr_sync(A));
    CHECK(tmk_amr_destroy(A));
    // a new layout of the finest level, data from the old layout and the next coarser level
    int lnx = nx, lny = ny;
    for (int l = 1; l < nlev; ++l) { lnx *= r; lny *= r; }
    std::vector<tmk_box> nb = finer_boxes(lnx / r, lny / r, h.boxes[nlev - 2], r);
    if (!nb.empty()) {
        tmk_level* fresh = nullptr;
        CHECK(tmk_level_create(h.ctx[nlev - 1], (int)nb.size(), nb.data(), &fresh));
        CHECK(tmk_level_regrid(fresh, h.lev[nlev - 1], h.lev[nlev - 2], r));
        CHECK(tmk_level_destroy(h.lev[nlev - 1]));
        h.lev[nlev - 1] = fresh;
        h.boxes[nlev - 1] = nb;
    }
    Hier g = make_hierarchy(nx, ny, nlev, r, precision, fault);
    if ((int)g.lev.size() == nlev) {
        tmk_amr* AB = nullptr;
        CHECK(tmk_amr_batch_create(nlev, r, &AB));
        CHECK(tmk_amr_add(AB, h.lev.data(), 0.0, 1e-3));
        CHECK(tmk_amr_add(AB, g.lev.data(), 0.0, 1e-3));
        CHECK(tmk_amr_advance(AB, 3));
        int code = 0;
        CHECK(tmk_amr_member_error(AB, h.lev[0], &code));
        CHECK(tmk_amr_remove(AB, g.lev[0]));
        CHECK(tmk_amr_advance(AB, 2));
        CHECK(tmk_amr_destroy(AB));
    }
    free_hierarchy(g);
    free_hierarchy(h);
}

// a batch of single-level scenarios joining and leaving
static void batch_case(int precision)
{
    tmk_batch* B = nullptr;
    CHECK(tmk_batch_create(&B));
    std::vector<tmk_ctx*> cs;
    std::vector<tmk_level*> ls;
    for (int k = 0; k < 4; ++k) {
        const int nx = uni(32, 96), ny = uni(16, 48);
        const tmk_params p = params(nx, ny, uni(0, 1), uni(0, 1), TMK_PRED_PPM, 25.0);
        tmk_ctx* c = nullptr;
        CHECK(tmk_ctx_create(&c, &p, precision));
        const std::vector<tmk_box> boxes = partition(nx, ny, uni(0, 3), false);
        tmk_level* L = nullptr;
        CHECK(tmk_level_create(c, (int)boxes.size(), boxes.data(), &L));
        CHECK(tmk_batch_add(B, L, 0.0, 1e-3));
        CHECK(tmk_batch_run(B, 2));
        cs.push_back(c);
        ls.push_back(L);
    }
    CHECK(tmk_batch_remove(B, ls[1]));
    CHECK(tmk_batch_run(B, 3));
    CHECK(tmk_batch_sync(B));
    CHECK(tmk_batch_destroy(B));
    for (tmk_level* L : ls) CHECK(tmk_level_destroy(L));
    for (tmk_ctx* c : cs) CHECK(tmk_ctx_destroy(c));
}

int main(int argc, char** argv)
{
    const int rounds = argc > 1 ? std::atoi(argv[1]) : 20;
    for (int k = 0; k < rounds; ++k) {
        const int precision = k % 2 ? TMK_PREC_FP32_BULK : TMK_PREC_FP64;
        level_case(precision);
        patch_case(precision);
        hierarchy_case(precision);
        if (k % 4 == 0) batch_case(precision);
        // every object is gone: the level-array cache holds idle blocks only (its blocks stay
        // reachable, so a level array that is never given back would escape LeakSanitizer)
        size_t cached = 0, live = 0, limit = 0;
        CHECK(tmk_cache_info(&cached, &live, &limit));
        if (live != 0) {
            std::fprintf(stderr, "round %d: %zu bytes of level arrays still in use\n", k, live);
            return 1;
        }
        if (k == rounds / 2) CHECK(tmk_cache_set_limit(1 << 20));
    }
    CHECK(tmk_cache_trim());
    std::printf("list builders: %d rounds clean\n", rounds);
    std::fflush(stdout);
    return 0;
}
