// C++ drop-in test: the plugin classes of LinElastHIP.H driven the way the
// reference drives LinElastPatchGodunov / LinElastLevelGodunov
// (LinElastLevelGodunov.cpp:273-324), checked against the CPU oracle.
#include <cmath>
#include <cstdio>
#include <random>

#include "LinElastHIP.H"
#include "tmo.h"

static void set_params(tmk_params& p, tmo_params& q, int nx, int ny)
{
    const double cs = 3464.0, cp = 6000.0, G = 2.67e-3 * cs * cs, dx = 25.0;
    p = tmk_params{};
    p.cs = cs; p.cp = cp; p.G = G;
    p.sxx0 = p.syy0 = p.szz0 = -120.0; p.sxy0 = 70.0;
    p.bc[0] = 0; p.bc[1] = 0; p.bc[2] = 0; p.bc[3] = 1;
    p.fault_kind = TMK_FAULT_RS;
    p.r0 = 1500.0; p.nsig = 25.0; p.rs_a = 0.014; p.rs_b = 0.018; p.rs_V0 = 1e-6; p.rs_f0 = 0.6; p.rs_L = 0.02;
    p.rs_fw = 0.13; p.rs_Vw = 0.17; p.rs_fexp = -1.0; p.x0 = nx * dx / 2; p.rs_rupture_thr = 1e-3;
    p.ramp_w = 1e40; p.rs_psi0 = 0.75; p.state_law = 0;
    p.normal_pred = TMK_PRED_PPM; p.fourth_order = 1; p.prim_limiting = 1; p.char_limiting = 0;
    p.dx = dx;
    p.domain.lo[0] = 0; p.domain.lo[1] = 0; p.domain.hi[0] = nx - 1; p.domain.hi[1] = ny - 1;
    q = tmo_params{};
    tmo_set_material(&q, cs, cp, G, p.sxx0, p.syy0, p.szz0, p.sxy0, 0.0, 0.0);
    for (int k = 0; k < 4; ++k) q.bc[k] = p.bc[k];
    q.fault_kind = 0; q.r0 = p.r0; q.nsig = p.nsig; q.rs_a = p.rs_a; q.rs_b = p.rs_b; q.rs_V0 = p.rs_V0;
    q.rs_f0 = p.rs_f0; q.rs_L = p.rs_L; q.rs_fw = p.rs_fw; q.rs_Vw = p.rs_Vw; q.rs_fExp = p.rs_fexp; q.x0 = p.x0;
    q.rs_rupture_thr = p.rs_rupture_thr; q.ramp_w = p.ramp_w; q.rs_psi0 = p.rs_psi0; q.state_law = 0;
    q.normal_pred = 2; q.fourth_order = 1; q.prim_limiting = 1; q.char_limiting = 0; q.dx = dx;
    q.domain.lo[0] = 0; q.domain.lo[1] = 0; q.domain.hi[0] = nx - 1; q.domain.hi[1] = ny - 1;
}

static double relerr(const double* a, const double* b, size_t n)
{
    double m = 0.0, s = 0.0;
    for (size_t k = 0; k < n; ++k) { m = std::fmax(m, std::fabs(a[k] - b[k])); s = std::fmax(s, std::fabs(b[k])); }
    return m / (s > 0 ? s : 1.0);
}

int main()
{
    const int nx = 48, ny = 32;
    tmk_params p;
    tmo_params q;
    set_params(p, q, nx, ny);
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U1(-1.0, 1.0);

    // --- patch level: the reference's per-box call sequence ---
    tmk::RSIBC ibc(p);
    tmk::LinElastPhysics phys(&ibc);
    tmk::LinElastPatchGodunov patch;
    const tmk::Box domain(0, 0, nx - 1, ny - 1);
    patch.define(domain, p.dx, &phys);
    const tmk::Box box(8, 0, 31, 19);
    tmk::FArrayBox U(tmk::grow(box, 4), 22), S;
    tmk::FArrayBox B(tmk::Box(box.lo[0] - 4, 0, box.hi[0] + 4, 0), 11);
    for (int c = 0; c < 22; ++c)
        for (int j = U.box().lo[1]; j <= U.box().hi[1]; ++j)
            for (int i = U.box().lo[0]; i <= U.box().hi[0]; ++i)
                U(i, j, c) = q.background[c] + ((c % 11 == 2 || c % 11 >= 7) ? 0.0 : (c % 11 < 3 ? 1.0 : 5.0) * U1(rng));
    B.setVal(0.0);
    for (int i = B.box().lo[0]; i <= B.box().hi[0]; ++i) { B(i, 0, 10) = 0.75 + 0.01 * U1(rng); B(i, 0, 9) = 1e10; }
    tmk::FArrayBox Ur = U, Br = B;
    tmk::FluxBox F;
    double mws = 0.0;
    const double time = 0.5, dt = 0.9 * 25.0 / 6000.0;
    patch.setCurrentTime(time);
    patch.setCurrentBox(box);
    ibc.setBdryData(&B);
    patch.updateState(U, F, mws, S, dt, box);
    tmo_fab uf{Ur.dataPtr(), {{Ur.box().lo[0], Ur.box().lo[1]}, {Ur.box().hi[0], Ur.box().hi[1]}}, 22};
    tmo_fab bf{Br.dataPtr(), {{Br.box().lo[0], 0}, {Br.box().hi[0], 0}}, 11};
    tmo_box vb{{box.lo[0], box.lo[1]}, {box.hi[0], box.hi[1]}};
    if (tmo_update_state(&uf, &bf, nullptr, nullptr, &q, vb, time, dt) != 0) { puts("oracle failed"); return 2; }
    const size_t nU = (size_t)U.box().size(0) * U.box().size(1) * 22;
    const double eu = relerr(U.dataPtr(), Ur.dataPtr(), nU), eb = relerr(B.dataPtr(), Br.dataPtr(), 11 * B.box().size(0));
    printf("patch: state rel err %.3e, fault %.3e, max wave speed %.1f\n", eu, eb, mws);
    if (!(eu <= 1e-12 && eb <= 1e-12 && mws == 6000.0)) return 1;

    // --- plastic pass: LinElastPatchGodunov::plasticUpdate (LinElastPatchGodunov.cpp:161-170) on a
    // box where about half the cells yield (Drucker-Prager with 1 MPa cohesion) ---
    {
        tmk_params pp = p;
        tmo_params qp = q;
        pp.use_plasticity = qp.use_plasticity = 1;
        pp.p_mu = qp.p_mu = 0.6;
        pp.p_beta = qp.p_beta = 0.3;
        pp.p_eta = qp.p_eta = 0.03 * pp.G;
        pp.p_b = qp.p_b = 1.0;
        pp.p_h = qp.p_h = 0.0;
        tmk::RSIBC pibc(pp);
        tmk::LinElastPhysics pphys(&pibc);
        tmk::LinElastPatchGodunov ppatch;
        ppatch.define(domain, pp.dx, &pphys);
        const tmk::Box pbox(4, 6, 35, 27);
        tmk::FArrayBox P(tmk::grow(pbox, 4), 22);
        for (int c = 0; c < 22; ++c)
            for (int j = P.box().lo[1]; j <= P.box().hi[1]; ++j)
                for (int i = P.box().lo[0]; i <= P.box().hi[0]; ++i)
                    P(i, j, c) = q.background[c] + (c % 11 == 6 ? 150.0 : (c % 11 >= 3 && c % 11 <= 5) ? 40.0 : 0.0) * U1(rng);
        tmk::FArrayBox Pr = P;
        ppatch.plasticUpdate(P, dt, pbox);
        tmo_fab pf{Pr.dataPtr(), {{Pr.box().lo[0], Pr.box().lo[1]}, {Pr.box().hi[0], Pr.box().hi[1]}}, 22};
        tmo_plastic_correction(&pf, &qp, dt, tmo_box{{pbox.lo[0], pbox.lo[1]}, {pbox.hi[0], pbox.hi[1]}});
        size_t yielded = 0, differ = 0;
        const size_t nP = (size_t)P.box().size(0) * P.box().size(1);
        for (size_t k = 0; k < nP * 22; ++k) differ += P.dataPtr()[k] != Pr.dataPtr()[k];
        for (size_t k = 0; k < nP; ++k) yielded += Pr.dataPtr()[k + 9 * nP] > 0.0;
        printf("plastic: %zu of %d cells yield, %zu values differ from the oracle\n", yielded,
               pbox.size(0) * pbox.size(1), differ);
        if (differ != 0 || yielded < (size_t)(pbox.size(0) * pbox.size(1)) / 4) return 1;
    }

    // --- level: LinElastLevelGodunov::step over a 2x2 box layout ---
    std::vector<tmk::Box> grids = {tmk::Box(0, 0, 23, 15), tmk::Box(24, 0, 47, 15), tmk::Box(0, 16, 23, 31),
                                   tmk::Box(24, 16, 47, 31)};
    tmk::LinElastLevelGodunov level;
    level.define(grids, &phys);
    const int g = 4, LX = nx + 2 * g, LY = ny + 2 * g;
    std::vector<double> Lu((size_t)LX * LY * 22), Lb((size_t)LX * 11, 0.0);
    for (int c = 0; c < 22; ++c)
        for (size_t k = 0; k < (size_t)LX * LY; ++k)
            Lu[k + c * (size_t)LX * LY] = q.background[c] + ((c % 11 == 2 || c % 11 >= 7) ? 0.0 : U1(rng));
    for (int i = 0; i < LX; ++i) { Lb[i + 10 * LX] = 0.75; Lb[i + 9 * LX] = 1e10; }
    for (int k = 0; k < 4; ++k) {
        const tmk::Box& v = grids[k];
        tmk::FArrayBox u(tmk::grow(v, 4), 22), b(tmk::Box(v.lo[0] - 4, 0, v.hi[0] + 4, 0), 11);
        for (int c = 0; c < 22; ++c)
            for (int j = u.box().lo[1]; j <= u.box().hi[1]; ++j)
                for (int i = u.box().lo[0]; i <= u.box().hi[0]; ++i)
                    u(i, j, c) = Lu[(size_t)(i + g) + (size_t)(j + g) * LX + (size_t)c * LX * LY];
        for (int c = 0; c < 11; ++c)
            for (int i = b.box().lo[0]; i <= b.box().hi[0]; ++i) b(i, 0, c) = Lb[(size_t)(i + g) + (size_t)c * LX];
        level.setData(k, u, v.lo[1] == 0 ? &b : nullptr);
    }
    double t = 0.0;
    for (int s = 0; s < 4; ++s) { level.step(t, dt); t = t + dt; }
    std::vector<double> Ou = Lu, Ob = Lb;
    if (tmo_level_steps(Ou.data(), Ob.data(), &q, nx, ny, 0.0, dt, 4, 1) != 0) { puts("oracle level failed"); return 2; }
    double worst = 0.0;
    for (int k = 0; k < 4; ++k) {
        const tmk::Box& v = grids[k];
        tmk::FArrayBox u(tmk::grow(v, 4), 22), b(tmk::Box(v.lo[0] - 4, 0, v.hi[0] + 4, 0), 11);
        level.getData(k, u, v.lo[1] == 0 ? &b : nullptr);
        for (int c = 0; c < 22; ++c) {
            double m = 0.0, sc = 0.0;
            for (int j = v.lo[1]; j <= v.hi[1]; ++j)
                for (int i = v.lo[0]; i <= v.hi[0]; ++i) {
                    const double o = Ou[(size_t)(i + g) + (size_t)(j + g) * LX + (size_t)c * LX * LY];
                    m = std::fmax(m, std::fabs(u(i, j, c) - o));
                    sc = std::fmax(sc, std::fabs(o));
                }
            worst = std::fmax(worst, m / (sc > 0 ? sc : 1.0));
        }
    }
    printf("level: 4 steps, 2x2 boxes, state rel err %.3e\n", worst);
    return worst <= 1e-12 ? 0 : 1;
}
