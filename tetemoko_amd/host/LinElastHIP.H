// LinElastHIP.H -- C++ host mirror of tetemoko's plugin surface over the C ABI.
//
// The class and method names, argument meanings and call order follow the
// reference so a Chombo host can swap its Fortran-backed classes for these:
//   LinElastPhysics           (srcLinElast/LinElastPhysics.H:44-265)
//   RSIBC / SWIBC             (ibcLinElast/RSIBC.H, ibcLinElast/SWIBC.H; LEPhysIBC.H:34-111)
//   LinElastPatchGodunov      (srcLinElast/LinElastPatchGodunov.H:78-122)
//   LinElastLevelGodunov      (srcLinElast/LinElastLevelGodunov.H:91-114)
// Box and FArrayBox here are minimal stand-ins with Chombo's storage convention
// (inclusive index box, Fortran order); a Chombo build passes its own
// FArrayBox::dataPtr()/box() to the same C ABI calls (include/tetemoko_hip.h).
// Errors the reference turns into MayDay::Error / Fortran stop are thrown as
// tmk::Error carrying the C ABI code.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "tetemoko_hip.h"

namespace tmk {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& where) : std::runtime_error(where + ": " + tmk_strerror(c)), code(c) {}
};

inline void check(int code, const char* where)
{
    if (code != TMK_OK) throw Error(code, where);
}

// Chombo-like cell-centred index box (2-D), inclusive
struct Box {
    int lo[2] = {0, 0}, hi[2] = {-1, -1};
    Box() = default;
    Box(int l0, int l1, int h0, int h1) { lo[0] = l0; lo[1] = l1; hi[0] = h0; hi[1] = h1; }
    int smallEnd(int d) const { return lo[d]; }
    int bigEnd(int d) const { return hi[d]; }
    int size(int d) const { return hi[d] - lo[d] + 1; }
    bool isEmpty() const { return hi[0] < lo[0] || hi[1] < lo[1]; }
    Box& grow(int n) { lo[0] -= n; lo[1] -= n; hi[0] += n; hi[1] += n; return *this; }
    tmk_box tmk() const { tmk_box b; b.lo[0] = lo[0]; b.lo[1] = lo[1]; b.hi[0] = hi[0]; b.hi[1] = hi[1]; return b; }
};

inline Box grow(Box b, int n) { return b.grow(n); }

// FArrayBox stand-in: data over an index box, x fastest, then y, then component
class FArrayBox {
  public:
    FArrayBox() = default;
    FArrayBox(const Box& b, int nc) { define(b, nc); }
    void define(const Box& b, int nc)
    {
        m_box = b;
        m_nc = nc;
        m_data.assign((size_t)b.size(0) * b.size(1) * nc, 0.0);
    }
    const Box& box() const { return m_box; }
    int nComp() const { return m_nc; }
    double* dataPtr() { return m_data.data(); }
    const double* dataPtr() const { return m_data.data(); }
    double& operator()(int i, int j, int c)
    {
        return m_data[(size_t)(i - m_box.lo[0]) + (size_t)(j - m_box.lo[1]) * m_box.size(0) +
                      (size_t)c * m_box.size(0) * m_box.size(1)];
    }
    void setVal(double v) { for (double& x : m_data) x = v; }

  private:
    Box m_box;
    int m_nc = 0;
    std::vector<double> m_data;
};

// FluxBox stand-in: one FArrayBox per direction on surroundingNodes(box, d)
class FluxBox {
  public:
    void resize(const Box& b, int nc)
    {
        m_f[0].define(Box(b.lo[0], b.lo[1], b.hi[0] + 1, b.hi[1]), nc);
        m_f[1].define(Box(b.lo[0], b.lo[1], b.hi[0], b.hi[1] + 1), nc);
    }
    FArrayBox& operator[](int d) { return m_f[d]; }

  private:
    FArrayBox m_f[2];
};

// Physical boundary conditions and fault law: LEPhysIBC (LEPhysIBC.H:34-111) with
// RSIBC (RSIBC.cpp:17-49) or SWIBC (SWIBC.cpp:15-50).  The constructor arguments of the
// reference and the COMMON-block setters fold into one tmk_params.
class LEPhysIBC {
  public:
    explicit LEPhysIBC(const tmk_params& p) : m_params(p) {}
    virtual ~LEPhysIBC() = default;
    // LEPhysIBC::setBdryData (LEPhysIBC.H:50): fault data of the current box (NULL off the fault)
    void setBdryData(FArrayBox* b) { m_bdry = b; }
    FArrayBox* bdryData() const { return m_bdry; }
    int numBdryVars() const { return m_params.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW; }
    const tmk_params& params() const { return m_params; }

  protected:
    tmk_params m_params;
    FArrayBox* m_bdry = nullptr;
};

class RSIBC : public LEPhysIBC {
  public:
    explicit RSIBC(tmk_params p) : LEPhysIBC((p.fault_kind = TMK_FAULT_RS, p)) {}
};

class SWIBC : public LEPhysIBC {
  public:
    explicit SWIBC(tmk_params p) : LEPhysIBC((p.fault_kind = TMK_FAULT_SW, p)) {}
};

// GodunovPhysics subclass (LinElastPhysics.H:44-265): owns the device context
class LinElastPhysics {
  public:
    explicit LinElastPhysics(LEPhysIBC* ibc, int precision = TMK_PREC_FP64) : m_ibc(ibc)
    {
        check(tmk_ctx_create(&m_ctx, &ibc->params(), precision), "LinElastPhysics");
    }
    ~LinElastPhysics() { tmk_ctx_destroy(m_ctx); }
    LinElastPhysics(const LinElastPhysics&) = delete;
    LinElastPhysics& operator=(const LinElastPhysics&) = delete;
    int numConserved() const { return TMK_NCOMP; }        // LinElastPhysics.cpp:69-74
    int numFluxes() const { return TMK_NCOMP; }           // :119-125
    LEPhysIBC* getPhysIBC() const { return m_ibc; }
    tmk_ctx* ctx() const { return m_ctx; }

  private:
    LEPhysIBC* m_ibc;
    tmk_ctx* m_ctx = nullptr;
};

// LinElastPatchGodunov (LinElastPatchGodunov.H:40-251)
class LinElastPatchGodunov {
  public:
    void define(const Box& domain, double dx, LinElastPhysics* phys)
    {
        m_domain = domain;
        m_dx = dx;
        m_phys = phys;
        m_defined = true;
    }
    void setCurrentTime(double t) { m_time = t; m_timeSet = true; }      // :86-90
    void setCurrentBox(const Box& b) { m_box = b; m_boxSet = true; }     // :93-99
    // updateState (LinElastPatchGodunov.cpp:101-129): U += dU on the box, F = fluxes,
    // maxWaveSpeed = max(cs, cp)
    void updateState(FArrayBox& U, FluxBox& F, double& maxWaveSpeed, const FArrayBox& /*S*/, double dt,
                     const Box& box)
    {
        if (!m_defined || !m_timeSet || !m_boxSet) throw Error(TMK_E_ARG, "LinElastPatchGodunov not defined");
        F.resize(box, m_phys->numFluxes());
        FArrayBox* B = m_phys->getPhysIBC()->bdryData();
        const bool onFault = box.smallEnd(1) == m_domain.smallEnd(1);
        check(tmk_patch_update_state(m_phys->ctx(), box.tmk(), U.dataPtr(), U.box().tmk(), U.nComp(),
                                     F[0].dataPtr(), F[1].dataPtr(), onFault && B ? B->dataPtr() : nullptr,
                                     onFault && B ? B->box().tmk() : box.tmk(), onFault && B ? B->nComp() : 0,
                                     m_time, dt, &maxWaveSpeed),
              "LinElastPatchGodunov::updateState");
    }
    // plasticUpdate (LinElastPatchGodunov.cpp:161-170)
    void plasticUpdate(FArrayBox& U, double dt, const Box& box)
    {
        check(tmk_patch_plastic_update(m_phys->ctx(), box.tmk(), U.dataPtr(), U.box().tmk(), U.nComp(), dt),
              "LinElastPatchGodunov::plasticUpdate");
    }

  private:
    Box m_domain, m_box;
    double m_dx = 0.0, m_time = 0.0;
    bool m_defined = false, m_timeSet = false, m_boxSet = false;
    LinElastPhysics* m_phys = nullptr;
};

// LinElastLevelGodunov (LinElastLevelGodunov.H:44-247), device resident:
// define() once per regrid, step() per level step.
class LinElastLevelGodunov {
  public:
    ~LinElastLevelGodunov() { if (m_level) tmk_level_destroy(m_level); }
    void define(const std::vector<Box>& grids, LinElastPhysics* phys)
    {
        std::vector<tmk_box> b;
        for (const Box& g : grids) b.push_back(g.tmk());
        m_grids = grids;
        check(tmk_level_create(phys->ctx(), (int)b.size(), b.data(), &m_level), "LinElastLevelGodunov::define");
    }
    // copy the level data in (the reference copies a_U into the 4-ghost scratch m_U, :211-224)
    void setData(int ibox, const FArrayBox& U, const FArrayBox* B)
    {
        check(tmk_level_upload(m_level, ibox, U.dataPtr(), B ? B->dataPtr() : nullptr), "LinElastLevelGodunov");
    }
    void getData(int ibox, FArrayBox& U, FArrayBox* B)
    {
        check(tmk_level_download(m_level, ibox, U.dataPtr(), B ? B->dataPtr() : nullptr), "LinElastLevelGodunov");
    }
    // step (LinElastLevelGodunov.cpp:174-398) + plasticUpdate (:643-667); returns the new dt
    double step(double time, double dt)
    {
        double dtNew = 0.0;
        check(tmk_level_step(m_level, time, dt, &dtNew), "LinElastLevelGodunov::step");
        check(tmk_level_sync(m_level), "LinElastLevelGodunov::step");
        return dtNew;
    }

  private:
    std::vector<Box> m_grids;
    tmk_level* m_level = nullptr;
};

}  // namespace tmk
