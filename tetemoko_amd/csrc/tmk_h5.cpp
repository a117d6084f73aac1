// Boundary plot files of the fault data in Chombo's AMR HDF5 layout (host code only).
//
// What goes into the file follows AMRLevelLinElast::writeCustomPlotFile and
// writeThisBdryLevel (srcLinElast/AMRLevelLinElast.cpp:1628-1756): the root header, then per
// level with fault boxes a group "level_<l>" with its header and the fault data over the
// level's fault boxes with output ghost IntVect::Unit.  How each item is laid out restates
// Chombo's HDF5 I/O (HDF5Handle creates group Chombo_global with SpaceDim and testReal;
// HDF5HeaderData writes scalar attributes, strings as fixed-length C strings of the string's
// length, boxes and IntVects as compound types; the LevelData write stores the layout as
// "boxes" and "Processors", the data as "<name>:datatype=0" with offsets "<name>:offsets=0",
// and "<name>_attributes" with comps, objectType, ghost and outputGhost, the output ghost
// clipped to the data's ghost vector).  Chombo is not in the reference tree, so that part is
// written from its published format.

#include "tetemoko_h5.h"

#include <hdf5.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

namespace {

constexpr int kGhostX = 4;  // x-ghosts of the fault data (AMRLevelLinElast.cpp:903-905)

// closes an HDF5 handle with the matching close call when it goes out of scope
struct Hid {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    Hid(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Hid() { if (id >= 0 && close) close(id); }
    Hid(const Hid&) = delete;
    Hid& operator=(const Hid&) = delete;
    bool ok() const { return id >= 0; }
};

struct Types {
    Hid box{H5Tcreate(H5T_COMPOUND, 4 * sizeof(int)), H5Tclose};
    Hid iv{H5Tcreate(H5T_COMPOUND, 2 * sizeof(int)), H5Tclose};
    bool ok = false;
    Types()
    {
        if (!box.ok() || !iv.ok()) return;
        ok = H5Tinsert(box.id, "lo_i", 0, H5T_NATIVE_INT) >= 0 &&
             H5Tinsert(box.id, "lo_j", sizeof(int), H5T_NATIVE_INT) >= 0 &&
             H5Tinsert(box.id, "hi_i", 2 * sizeof(int), H5T_NATIVE_INT) >= 0 &&
             H5Tinsert(box.id, "hi_j", 3 * sizeof(int), H5T_NATIVE_INT) >= 0 &&
             H5Tinsert(iv.id, "intvecti", 0, H5T_NATIVE_INT) >= 0 &&
             H5Tinsert(iv.id, "intvectj", sizeof(int), H5T_NATIVE_INT) >= 0;
    }
};

bool attr(hid_t loc, const char* name, hid_t type, const void* v)
{
    Hid sp(H5Screate(H5S_SCALAR), H5Sclose);
    if (!sp.ok()) return false;
    Hid a(H5Acreate2(loc, name, type, sp.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return a.ok() && H5Awrite(a.id, type, v) >= 0;
}

bool attr_int(hid_t loc, const char* name, int v) { return attr(loc, name, H5T_NATIVE_INT, &v); }
bool attr_real(hid_t loc, const char* name, double v) { return attr(loc, name, H5T_NATIVE_DOUBLE, &v); }

bool attr_str(hid_t loc, const char* name, const std::string& s)
{
    Hid t(H5Tcopy(H5T_C_S1), H5Tclose);
    if (!t.ok() || s.empty() || H5Tset_size(t.id, s.size()) < 0) return false;
    return attr(loc, name, t.id, s.data());
}

// 1-D dataset of n elements of type (an element may span several entries of v)
template <class T>
bool dataset(hid_t loc, const char* name, hid_t type, const std::vector<T>& v, size_t n_elem)
{
    const hsize_t n = n_elem;
    Hid sp(H5Screate_simple(1, &n, nullptr), H5Sclose);
    if (!sp.ok()) return false;
    Hid d(H5Dcreate2(loc, name, type, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    return d.ok() && H5Dwrite(d.id, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()) >= 0;
}

bool level_ok(const tmk_bdry_level& L)
{
    if (L.level < 0 || L.nbox < 0 || L.domain_nx <= 0) return false;
    if (L.nbox == 0) return true;
    if (!L.boxes || !L.data) return false;
    for (int b = 0; b < L.nbox; ++b) {
        const int* q = L.boxes + 4 * b;
        if (q[0] < 0 || q[2] < q[0] || q[2] >= L.domain_nx || q[1] != 0 || q[3] != 0 || !L.data[b]) return false;
    }
    return true;
}

// writeThisBdryLevel (AMRLevelLinElast.cpp:1713-1749) for one level with fault boxes
bool write_level(hid_t file, const Types& T, const tmk_bdry_level& L, int ncomp)
{
    char label[32];
    std::snprintf(label, sizeof label, "level_%d", L.level);
    Hid g(H5Gcreate2(file, label, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    if (!g.ok()) return false;
    const int dom[4] = {0, 0, L.domain_nx - 1, 0};  // bdryLo(domain, 1, 1)
    if (!attr_int(g.id, "ref_ratio", L.ref_ratio) || !attr_real(g.id, "dx", L.dx) ||
        !attr_real(g.id, "dt", L.dt) || !attr_real(g.id, "time", L.time) ||
        !attr(g.id, "prob_domain", T.box.id, dom))
        return false;

    // the layout's boxes in Chombo's order (sorted by the low corner, i first)
    std::vector<int> order(L.nbox);
    for (int b = 0; b < L.nbox; ++b) order[b] = b;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const int* p = L.boxes + 4 * a;
        const int* q = L.boxes + 4 * b;
        return p[0] != q[0] ? p[0] < q[0] : p[1] < q[1];
    });
    std::vector<int> boxes(4 * (size_t)L.nbox), procs(L.nbox, 0);
    std::vector<long long> offs(L.nbox + 1, 0);
    const int ogx = std::min(1, kGhostX);  // requested output ghost (1,1) clipped to (4,0)
    size_t total = 0;
    for (int k = 0; k < L.nbox; ++k) {
        const int* q = L.boxes + 4 * order[k];
        std::copy(q, q + 4, boxes.begin() + 4 * k);
        total += (size_t)ncomp * (q[2] - q[0] + 1 + 2 * ogx);
        offs[k + 1] = (long long)total;
    }
    std::vector<double> data;
    data.reserve(total);
    for (int k = 0; k < L.nbox; ++k) {
        const int* q = L.boxes + 4 * order[k];
        const int n = q[2] - q[0] + 1;
        const double* src = L.data[order[k]];
        for (int c = 0; c < ncomp; ++c) {
            const double* row = src + (size_t)c * (n + 2 * kGhostX) + (kGhostX - ogx);
            data.insert(data.end(), row, row + n + 2 * ogx);
        }
    }
    if (!dataset(g.id, "boxes", T.box.id, boxes, L.nbox) ||
        !dataset(g.id, "Processors", H5T_NATIVE_INT, procs, procs.size()) ||
        !dataset(g.id, "data:offsets=0", H5T_NATIVE_LLONG, offs, offs.size()) ||
        !dataset(g.id, "data:datatype=0", H5T_NATIVE_DOUBLE, data, data.size()))
        return false;
    Hid a(H5Gcreate2(g.id, "data_attributes", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
    const int ghost[2] = {kGhostX, 0}, og[2] = {ogx, 0};
    return a.ok() && attr_int(a.id, "comps", ncomp) && attr_str(a.id, "objectType", "FArrayBox") &&
           attr(a.id, "ghost", T.iv.id, ghost) && attr(a.id, "outputGhost", T.iv.id, og);
}

bool write_file(const char* path, int max_level, int finest_level, int step, double time, int ncomp,
                const char* const* names, int nlev, const tmk_bdry_level* levels)
{
    Types T;
    if (!T.ok) return false;
    Hid f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (!f.ok()) return false;
    {
        Hid g(H5Gcreate2(f.id, "Chombo_global", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        if (!g.ok() || !attr_int(g.id, "SpaceDim", 2) || !attr_real(g.id, "testReal", 0.0)) return false;
    }
    // writeCustomPlotFile's header (AMRLevelLinElast.cpp:1665-1685)
    Hid root(H5Gopen2(f.id, "/", H5P_DEFAULT), H5Gclose);
    if (!root.ok() || !attr_int(root.id, "max_level", max_level) ||
        !attr_int(root.id, "num_levels", finest_level + 1) || !attr_int(root.id, "iteration", step) ||
        !attr_real(root.id, "time", time) || !attr_int(root.id, "num_components", ncomp))
        return false;
    for (int c = 0; c < ncomp; ++c) {
        char key[32];
        std::snprintf(key, sizeof key, "component_%d", c);
        if (!attr_str(root.id, key, names[c])) return false;
    }
    for (int l = 0; l < nlev; ++l)
        if (levels[l].nbox > 0 && !write_level(f.id, T, levels[l], ncomp)) return false;
    return H5Fflush(f.id, H5F_SCOPE_LOCAL) >= 0;
}

}  // namespace

extern "C" int tmk_h5_write_bdry_plot(const char* path, int max_level, int finest_level, int step, double time,
                                      int ncomp, const char* const* names, int nlev,
                                      const tmk_bdry_level* levels)
{
    if (!path || ncomp <= 0 || !names || nlev < 0 || (nlev > 0 && !levels) || finest_level < 0 ||
        max_level < finest_level)
        return TMK_E_ARG;
    for (int c = 0; c < ncomp; ++c)
        if (!names[c] || !names[c][0]) return TMK_E_ARG;
    for (int l = 0; l < nlev; ++l)
        if (!level_ok(levels[l])) return TMK_E_ARG;
    // the library's own error printing is off for the call; failures come back as TMK_E_IO
    H5E_auto2_t fn = nullptr;
    void* data = nullptr;
    H5Eget_auto2(H5E_DEFAULT, &fn, &data);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    const bool ok = write_file(path, max_level, finest_level, step, time, ncomp, names, nlev, levels);
    H5Eset_auto2(H5E_DEFAULT, fn, data);
    return ok ? 0 : TMK_E_IO;
}

extern "C" int tmk_h5_bdry_plot_name(const char* prefix, int step, char* out, int cap)
{
    if (!prefix || !out || cap <= 0) return -1;
    const int n = std::snprintf(out, cap, "%sboundary.%06d.%dd.hdf5", prefix, step, 2);
    return (n < 0 || n >= cap) ? -1 : n;
}
