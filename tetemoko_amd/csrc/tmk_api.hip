// tmk_api.hip -- host side of libtetemoko_hip: contexts, device-resident levels,
// multi-scenario batches and the C ABI of include/tetemoko_hip.h.
//
// A "group" is the unit of one fused launch: a device array of scenario blocks
// (parameters + clock), the box descriptors of every member level for both
// buffer parities, the tile work list and the ghost-exchange plan.  A level
// stepping on its own is a one-member group; a batch concatenates the levels of
// many scenarios into one group so that their boxes share each launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>
#include <array>

#include "tetemoko_hip.h"

#include "tmk_kernels.hpp"
#include "tmk_interior.hpp"
#include "tmk_march.hpp"
#include "tmk_amr.hpp"

#ifndef TMK_DEFAULT_BY_F64
#define TMK_DEFAULT_BY_F64 8
#endif
// strip rows of the row-marching interior update for new contexts (measured on the 8192 x 4096
// and 4096 x 2048 rate-and-state levels: 32 rows is within 3 % of the best of 16 ... 256 for FP32
// bulk and FP64 on both)
#ifndef TMK_DEFAULT_MARCH_ROWS
#define TMK_DEFAULT_MARCH_ROWS 32
#endif

static void default_tile(int precision, int& bx, int& by);
static int scratch_device_ptr(void** out);

// the FP32-bulk update kernels are instantiated in tmk_f32.hip (built with FMA contraction)
namespace tmk {
#define TMK_DECLARE_EXTERN(T, NS, BX, BY, PK) \
    extern template __global__ void k_godunov_interior<T, NS, BX, BY, PK>(const KArgs<T>);
TMK_ALL_UPDATE_KERNELS(TMK_DECLARE_EXTERN, float)
extern template __global__ void k_godunov_march<float, PK_PPM_PRIM, true>(const KArgs<float>);
extern template __global__ void k_godunov_march<float, PK_PPM_PRIM, false>(const KArgs<float>);
extern template __global__ void k_godunov_bnd<float, 2, PK_RUNTIME>(const KArgs<float>);
extern template __global__ void k_godunov_bnd<float, 3, PK_RUNTIME>(const KArgs<float>);
extern template __global__ void k_godunov_bnd<float, 2, PK_PPM_PRIM>(const KArgs<float>);
extern template __global__ void k_godunov_bnd<float, 3, PK_PPM_PRIM>(const KArgs<float>);
}  // namespace tmk

using namespace tmk;

#define HIPCHK(x)                                                                           \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            fprintf(stderr, "tetemoko_hip: %s failed: %s\n", #x, hipGetErrorString(e_));   \
            return e_ == hipErrorOutOfMemory ? TMK_E_NOMEM : TMK_E_HIP;                     \
        }                                                                                   \
    } while (0)

namespace {

// grow-only device buffer owned by one object: freed when that object goes away
struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    int ensure(size_t bytes);
    void release();
};

// ---------------------------------------------------------------------------
// Block cache for level arrays.  A serving loop creates and destroys levels without pause, and
// hipMalloc / hipFree of their arrays cost milliseconds each while the device is busy (hipFree
// waits for it).  The island, fault-row and owner arrays of levels and the flux arrays and face
// masks of hierarchy members come from a per-device cache of blocks in size classes (at most 1/8
// above the request) and go back to it.  A block goes back only when no queued work can touch it:
// tmk_level_destroy refuses batch members and synchronises the level's stream, and a hierarchy
// member is released after its batch's stream has been synchronised.  Cached bytes count as free
// in tmk_mem_info; a failed hipMalloc empties the device's cache and tries once more.
struct BlockCache {
    std::mutex mu;
    std::map<std::pair<int, size_t>, std::vector<void*>> idle;  // (device, class bytes) -> blocks
    std::unordered_map<void*, std::pair<int, size_t>> live;      // block -> (device, class bytes)
    std::map<int, size_t> cached, in_use;                        // bytes per device
    size_t limit = (size_t)32 << 30;                             // cached bytes per device, at most
};

static BlockCache& block_cache()
{
    static BlockCache* c = new BlockCache;  // never destroyed: blocks stay valid until exit
    return *c;
}

static size_t size_class(size_t n)
{
    if (n < 256) return 256;
    const size_t g = std::max<size_t>(256, ((size_t)1 << (63 - __builtin_clzll(n))) >> 3);
    return (n + g - 1) / g * g;
}

// frees every idle block of device dev (the caller holds no lock)
static void cache_trim_device(int dev, size_t keep)
{
    BlockCache& C = block_cache();
    std::vector<void*> drop;
    {
        std::lock_guard<std::mutex> lk(C.mu);
        for (auto it = C.idle.begin(); it != C.idle.end() && C.cached[dev] > keep;) {
            if (it->first.first != dev) { ++it; continue; }
            while (!it->second.empty() && C.cached[dev] > keep) {
                drop.push_back(it->second.back());
                it->second.pop_back();
                C.cached[dev] -= it->first.second;
            }
            it = it->second.empty() ? C.idle.erase(it) : std::next(it);
        }
    }
    for (void* p : drop) (void)hipFree(p);
}

static hipError_t dev_alloc(void** p, size_t n)
{
    BlockCache& C = block_cache();
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const size_t cls = size_class(n);
    {
        std::lock_guard<std::mutex> lk(C.mu);
        auto it = C.idle.find({dev, cls});
        if (it != C.idle.end() && !it->second.empty()) {
            *p = it->second.back();
            it->second.pop_back();
            C.cached[dev] -= cls;
            C.in_use[dev] += cls;
            C.live[*p] = {dev, cls};
            return hipSuccess;
        }
    }
    e = hipMalloc(p, cls);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        cache_trim_device(dev, 0);
        e = hipMalloc(p, cls);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    std::lock_guard<std::mutex> lk(C.mu);
    C.live[*p] = {dev, cls};
    C.in_use[dev] += cls;
    return hipSuccess;
}

template <typename X>
static hipError_t dev_alloc(X** p, size_t n)
{
    void* q = nullptr;
    const hipError_t e = dev_alloc(&q, n);
    *p = (X*)q;
    return e;
}

static void dev_free(void* p)
{
    if (!p) return;
    BlockCache& C = block_cache();
    {
        std::lock_guard<std::mutex> lk(C.mu);
        auto it = C.live.find(p);
        if (it != C.live.end()) {
            const int dev = it->second.first;
            const size_t cls = it->second.second;
            C.live.erase(it);
            C.in_use[dev] -= cls;
            if (C.cached[dev] + cls <= C.limit) {
                C.idle[{dev, cls}].push_back(p);
                C.cached[dev] += cls;
                return;
            }
        }
    }
    (void)hipFree(p);
}

// A DevBuf's block comes from the block cache; it grows or goes back only when no queued work
// reads it (every owner synchronises the streams that use it first).
int DevBuf::ensure(size_t bytes)
{
    if (bytes <= n) return TMK_OK;
    release();
    if (dev_alloc(&p, bytes) != hipSuccess) return TMK_E_NOMEM;
    n = bytes;
    return TMK_OK;
}

void DevBuf::release()
{
    dev_free(p);
    p = nullptr;
    n = 0;
}

// A copy ordered on stream st, complete on return.  hipMemcpy runs on the null stream, which
// waits for every stream of the device: a serving loop's host work would stall behind other
// scenarios' running launches.
static hipError_t copy_on(hipStream_t st, void* dst, const void* src, size_t n, hipMemcpyKind kind)
{
    if (!n) return hipSuccess;
    hipError_t e = hipMemcpyAsync(dst, src, n, kind, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
}

inline bool box_contains(const tmk_box& o, const tmk_box& b)
{
    return b.lo[0] >= o.lo[0] && b.lo[1] >= o.lo[1] && b.hi[0] <= o.hi[0] && b.hi[1] <= o.hi[1];
}
inline tmk_box box_and(tmk_box a, const tmk_box& b)
{
    for (int d = 0; d < 2; ++d) {
        a.lo[d] = std::max(a.lo[d], b.lo[d]);
        a.hi[d] = std::min(a.hi[d], b.hi[d]);
    }
    return a;
}
inline tmk_box box_grow(tmk_box a, int n)
{
    a.lo[0] -= n; a.lo[1] -= n; a.hi[0] += n; a.hi[1] += n;
    return a;
}
inline bool box_empty(const tmk_box& a) { return a.hi[0] < a.lo[0] || a.hi[1] < a.lo[1]; }

void make_tiles(const tmk_box& v, int patch, int tx, int ty, std::vector<Tile>& out)
{
    for (int y0 = v.lo[1]; y0 <= v.hi[1]; y0 += ty)
        for (int x0 = v.lo[0]; x0 <= v.hi[0]; x0 += tx) {
            Tile t;
            t.patch = patch;
            t.x0 = x0;
            t.y0 = y0;
            t.x1 = std::min(x0 + tx - 1, v.hi[0]);
            t.y1 = std::min(y0 + ty - 1, v.hi[1]);
            out.push_back(t);
        }
}

// Workgroup g runs on XCD g % 8.  Give each XCD a contiguous band of the tile
// list so that halos shared by neighbouring tiles stay in that XCD's L2.
void xcd_order(std::vector<Tile>& t)
{
    const size_t nx = 8, n = t.size();
    if (n < 64) return;
    const size_t per = (n + nx - 1) / nx;
    std::vector<Tile> o;
    o.reserve(n);
    for (size_t g = 0; g < per * nx; ++g) {
        const size_t src = (g % nx) * per + g / nx;
        if (src < n) o.push_back(t[src]);
    }
    t.swap(o);
}

// a tile is interior when its stencil (+-3 cells) lies inside the domain (the same rule as the
// kernels' tile_interior); partial tiles at box edges qualify
bool tile_interior_host(const Tile& t, const DevParams& P, int bx, int by)
{
    return t.x1 - t.x0 + 1 <= bx - 2 && t.y1 - t.y0 + 1 <= by - 2 && t.x0 - 3 >= P.dlo[0] && t.x1 + 3 <= P.dhi[0] &&
           t.y0 - 3 >= P.dlo[1] && t.y1 + 3 <= P.dhi[1];
}

// The cells of rectangle r within 3 of the domain's edges, whose stencils need the boundary
// conditions, form bands at most 3 cells thick: the rows below and above r's core (full width)
// and the columns left and right of it.  band_split passes them to put as boundary pieces --
// (BND_BX - 2) x (BND_BY - 2) along a row band, the transpose along a column band -- and returns
// the core, the cells whose +-3 stencil lies inside the domain (empty when there are none).
template <typename Put>
Tile band_split(const Tile& r, const DevParams& P, Put put)
{
    Tile c = r;
    c.x0 = std::max(r.x0, P.dlo[0] + 3);
    c.x1 = std::min(r.x1, P.dhi[0] - 3);
    c.y0 = std::max(r.y0, P.dlo[1] + 3);
    c.y1 = std::min(r.y1, P.dhi[1] - 3);
    auto pieces = [&](int x0, int y0, int x1, int y1, bool col) {
        const int pw = col ? BND_BY - 2 : BND_BX - 2, ph = col ? BND_BX - 2 : BND_BY - 2;
        for (int py = y0; py <= y1; py += ph)
            for (int px = x0; px <= x1; px += pw) {
                Tile t;
                t.patch = r.patch;
                t.x0 = px;
                t.y0 = py;
                t.x1 = std::min(px + pw - 1, x1);
                t.y1 = std::min(py + ph - 1, y1);
                put(t);
            }
    };
    if (c.x0 > c.x1 || c.y0 > c.y1) {
        // no core: r lies in a band (or the domain is under 7 cells across)
        pieces(r.x0, r.y0, r.x1, r.y1, r.x1 - r.x0 + 1 <= BND_BY - 2 && r.y1 - r.y0 + 1 > BND_BY - 2);
        c.x1 = c.x0 - 1;
        return c;
    }
    pieces(r.x0, r.y0, r.x1, c.y0 - 1, false);  // band below the core
    pieces(r.x0, c.y1 + 1, r.x1, r.y1, false);  // above
    pieces(r.x0, c.y0, c.x0 - 1, c.y1, true);   // left of it
    pieces(c.x1 + 1, c.y0, r.x1, c.y1, true);   // right of it
    return c;
}

inline bool tile_empty(const Tile& t) { return t.x0 > t.x1 || t.y0 > t.y1; }

// interior tiles first (XCD order), then the boundary pieces, fault-row pieces first (they
// carry the fault solve, the longest workgroups); a tile that is not interior contributes its
// core as an interior tile and its bands as pieces.  P(t): the parameters of tile t's scenario.
// Returns the interior count.
template <typename F, typename G>
int partition_tiles(std::vector<Tile>& t, F is_interior, G params, bool xcd = true)
{
    std::vector<Tile> in, bf, bo;
    for (const Tile& x : t) {
        if (is_interior(x)) {
            in.push_back(x);
            continue;
        }
        const DevParams& P = params(x);
        const Tile c = band_split(x, P, [&](const Tile& p) { (p.y0 == P.dlo[1] ? bf : bo).push_back(p); });
        if (!tile_empty(c)) in.push_back(c);
    }
    if (xcd) xcd_order(in);
    const int nint = (int)in.size();
    t.swap(in);
    t.insert(t.end(), bf.begin(), bf.end());
    t.insert(t.end(), bo.begin(), bo.end());
    return nint;
}

// Row-marching partition of one rectangle of an island (r.patch: its descriptor): the core in
// strips of MARCH_COLS x rows for k_godunov_march, the bands in boundary pieces (fault-row
// pieces apart).
void march_rect(const Tile& r, const DevParams& P, int rows, std::vector<Tile>& strips, std::vector<Tile>& bf,
                std::vector<Tile>& bo)
{
    const Tile c = band_split(r, P, [&](const Tile& p) { (p.y0 == P.dlo[1] ? bf : bo).push_back(p); });
    if (tile_empty(c)) return;
    for (int y = c.y0; y <= c.y1; y += rows)
        for (int x = c.x0; x <= c.x1; x += MARCH_COLS) {
            Tile t;
            t.patch = r.patch;
            t.x0 = x;
            t.y0 = y;
            t.x1 = std::min(x + MARCH_COLS - 1, c.x1);
            t.y1 = std::min(y + rows - 1, c.y1);
            strips.push_back(t);
        }
}

// state planes that sweeps read with a stencil (double-buffered); the others
// (szz, lam, gam and, in mode II, the anti-plane planes) are updated in place
bool is_stencil(int c, int ns)
{
    const int k = c % 11;
    if (k == 0 || k == 1 || k == 3 || k == 4 || k == 6) return true;
    if (ns == 3 && (k == 2 || k == 7 || k == 8)) return true;
    return false;
}

int stencil_mask(int ns)
{
    int m = 0;
    for (int c = 0; c < TMK_NCOMP; ++c)
        if (is_stencil(c, ns)) m |= 1 << c;
    return m;
}

}  // namespace

struct Group {
    size_t tsize = 8;
    int ns = 2, bx = 32, by = 16;
    int march = 0;                 // > 0: interior strips of this many rows run k_godunov_march
    bool any_plast = true;         // some member has plasticity (the march variant with lam, gam)
    int xcd = 1;                   // interior tiles / strips in XCD bands (the first member's choice)
    int pk = PK_RUNTIME;  // predictor kind the kernel is specialised for (all members agree)
    Scen* scen_dev = nullptr;
    int nscen = 0;
    void* patches_dev[2] = {nullptr, nullptr};
    Tile* tiles_dev = nullptr;     // interior tiles first, then the boundary tiles
    int ntiles = 0;
    int nint = 0;                  // interior tiles (k_godunov_interior); the rest run k_godunov
    long long cells_int = 0, cells_bnd = 0;
    hipStream_t side = nullptr;    // boundary tiles run here, concurrently with the interior launch
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    CopyRect* rects_dev = nullptr;
    int nrects = 0;
    CopyRect* brects_dev = nullptr;
    int nbrects = 0;
    int nb = 11;
    int plastic = 1;
    const StationDev* stations = nullptr;  // fault stations recorded after each step (level mode)
    int g = 0;  // parity of the next launch
    hipGraphExec_t gexec = nullptr;
    // grow-only backing of the device tables above: a rebuild (membership change)
    // rewrites them in place, so no device allocation or free per rebuild
    DevBuf mem_scen, mem_pt[2], mem_tiles, mem_rects, mem_brects;
    bool live() const { return nscen > 0; }
    // forget the layout; keeps the side stream, events and device tables for reuse.
    // The caller has synchronised the stream the group ran on.
    void reset()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        scen_dev = nullptr; tiles_dev = nullptr; rects_dev = nullptr; brects_dev = nullptr;
        patches_dev[0] = patches_dev[1] = nullptr;
        nscen = ntiles = nint = nrects = nbrects = 0;
    }
    void release()
    {
        reset();
        if (side) (void)hipStreamSynchronize(side);
        mem_scen.release(); mem_pt[0].release(); mem_pt[1].release();
        mem_tiles.release(); mem_rects.release(); mem_brects.release();
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (side) (void)hipStreamDestroy(side);
        side = nullptr; ev_fork = ev_join = nullptr;
    }
};

// Every stream the library creates is non-blocking: a synchronous copy (hipMemcpy, the null
// stream) does not wait for other scenarios' work in flight.  Every copy that reads device
// results synchronises the stream that produced them first.
struct tmk_ctx {
    tmk_params hp;
    DevParams P;
    int precision = TMK_PREC_FP64;
    hipStream_t stream = nullptr;
    std::vector<double*> face_dev;
    DevBuf bufs[8];  // grow-only device buffers of the drop-in path, uploads and tags
    int bx = 32, by = 16;
    bool antiplane = true;
    int march = TMK_DEFAULT_MARCH_ROWS;  // strip rows of the row-marching interior update (0: tile kernels)
    void* noalloc_scratch = nullptr;  // this device's scratch array (tmk_patch_update_state_noalloc)
};

struct tmk_batch;
struct tmk_stations;

// Boxes whose ghost regions reach each other share one set of arrays over their union's
// bounding box (grown by the ghost width): an island.  Neighbouring boxes then read each
// other's cells in place (no ghost exchange inside an island), and the update tiles cover
// merged rectangles of the union instead of each box.  An island whose bounding box would
// be more than twice the boxes' own ghosted area is split into one island per box.
struct Island {
    std::vector<int> boxes;
    tmk_box bbox{}, abox{};      // union bounding box; array box = bbox grown by the ghosts
    int has_b = 0;               // a box of the island lies on the fault
    int blo0 = 0, bnx = 0;       // fault-data row: faces blo0 .. blo0 + bnx - 1
    void* u[2] = {nullptr, nullptr};  // stencil planes (ping-pong)
    void* w = nullptr;                // planes updated in place
    double* b[2] = {nullptr, nullptr};
    int* owner = nullptr;        // device: box owning each array cell, -1 for ghosts
    int* bowner = nullptr;       // device: box owning each fault face, -1 for ghosts
    size_t plane() const { return (size_t)(abox.hi[0] - abox.lo[0] + 1) * (abox.hi[1] - abox.lo[1] + 1); }
    int unx() const { return abox.hi[0] - abox.lo[0] + 1; }
};

struct tmk_level {
    tmk_ctx* ctx = nullptr;
    std::vector<tmk_box> boxes;
    std::vector<int> has_b;
    int ns = 2;
    size_t tsize = 8;
    int bx = 32, by = 16;         // update tile (the context's at creation; tmk_level_set_update_shape)
    int march = 0;                // strip rows of the row-marching update (0: tile kernels)
    int xcd = 1;                  // interior tiles / strips in XCD bands (0: row-major order)
    std::vector<Island> isl;
    std::vector<int> box_isl;     // island of each box
    int npatch = 0;               // descriptors: one per box, then one per island (the tiles')
    int cur = 0;                  // buffer parity holding the current state
    double time = 0.0, dt = 0.0;  // clock of this scenario
    std::vector<char> hpatch[2];  // Patch<T> descriptors for parity 0 / 1 (scen index 0)
    std::vector<Tile> tiles;
    std::vector<Tile> urects;     // the islands' merged rectangles (patch: the island descriptor)
    std::vector<CopyRect> rects, brects;
    Group own;
    bool own_built = false;
    bool dirty = true;            // uploaded data not yet exchanged
    tmk_batch* batch = nullptr;
    hipStream_t amr_stream = nullptr;  // set while the level is part of a hierarchy (tmk_amr)
    size_t bytes = 0;             // device bytes held
    tmk_stations* stations = nullptr;
};

struct tmk_stations {
    tmk_level* L = nullptr;
    StationDev dev{};
    DevBuf meta, rec, count;
};

struct tmk_batch {
    hipStream_t stream = nullptr;
    std::vector<tmk_level*> members;
    std::vector<int> phase;       // member parity at group parity 0
    Group grp;
    Group xgrp;                   // one-member group for the ghost exchange of a newly added member
    bool built = false;
    bool stale = false;           // membership changed since the group was built
};

// largest state plane the kernels address (32-bit byte offsets from the plane's base, and at
// most kDropOffset, the offset the row-marching kernel gives lanes whose accesses must be
// dropped); one FP64 level of planes that size (three copies of 22 planes) would exceed the
// device's memory anyway
static constexpr size_t kPlaneMaxBytes = kDropOffset;

static const char* k_errs[] = {"ok",
                               "NaN state variable psi (dPsi is nan)",
                               "invalid boundary type",
                               "invalid normal predictor or limiter combination",
                               "fault face without fault data",
                               "array box does not cover the update stencil",
                               "invalid argument",
                               "HIP runtime failure",
                               "unsupported option",
                               "device memory exhausted",
                               "station records dropped: buffer full before it was read",
                               "coarse-fine time interpolation weight outside [0, 1]",
                               "finer level not aligned to the refinement ratio or not properly nested",
                               "file output failed",
                               "box exceeds the fixed scratch of the allocation-free drop-in"};

extern "C" const char* tmk_strerror(int code)
{
    const int k = -code;
    if (k < 0 || k >= (int)(sizeof(k_errs) / sizeof(k_errs[0]))) return "unknown error";
    return k_errs[k];
}

extern "C" int tmk_device_ok(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    return strstr(prop.gcnArchName, "gfx950") != nullptr ? 1 : 0;
}

extern "C" int tmk_mem_info(size_t* free_bytes, size_t* total_bytes)
{
    if (!free_bytes || !total_bytes) return TMK_E_ARG;
    HIPCHK(hipMemGetInfo(free_bytes, total_bytes));
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    BlockCache& C = block_cache();
    std::lock_guard<std::mutex> lk(C.mu);
    *free_bytes += C.cached[dev];
    return TMK_OK;
}

extern "C" int tmk_cache_info(size_t* cached_bytes, size_t* live_bytes, size_t* limit_bytes)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    BlockCache& C = block_cache();
    std::lock_guard<std::mutex> lk(C.mu);
    if (cached_bytes) *cached_bytes = C.cached[dev];
    if (live_bytes) *live_bytes = C.in_use[dev];
    if (limit_bytes) *limit_bytes = C.limit;
    return TMK_OK;
}

extern "C" int tmk_cache_set_limit(size_t bytes)
{
    BlockCache& C = block_cache();
    std::vector<int> devs;
    {
        std::lock_guard<std::mutex> lk(C.mu);
        C.limit = bytes;
        for (const auto& kv : C.cached) devs.push_back(kv.first);
    }
    for (int d : devs) cache_trim_device(d, bytes);
    return TMK_OK;
}

extern "C" int tmk_cache_trim(void)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    cache_trim_device(dev, 0);
    return TMK_OK;
}

static void fill_dev_params(const tmk_params& h, DevParams& P)
{
    memset(&P, 0, sizeof P);
    // LINELASTSETF (LinElastPhysicsF.ChF:1165-1207)
    P.cs = h.cs; P.cp = h.cp; P.G = h.G;
    P.rho = h.G / (h.cs * h.cs);
    P.lam = (h.cp * h.cp) * P.rho - 2.0 * h.G;
    P.Zs = P.rho * h.cs;
    P.Zp = P.rho * h.cp;
    P.iZp = 1.0 / P.Zp;
    P.iZs = 1.0 / P.Zs;
    P.irho = 1.0 / P.rho;
    P.i2Zp = 0.5 * P.iZp;
    P.i2Zs = 0.5 * P.iZs;
    P.K = P.lam + 2.0 * h.G / 3.0;
    P.den = 2.0 * h.G + P.lam;
    P.lam2G = P.lam + 2.0 * h.G;
    P.frho = (float)P.rho; P.flam = (float)P.lam; P.fG = (float)P.G; P.fZp = (float)P.Zp; P.fZs = (float)P.Zs;
    P.fden = (float)P.den; P.flam2G = (float)P.lam2G; P.fiZp = (float)P.iZp; P.fiZs = (float)P.iZs;
    P.firho = (float)P.irho; P.fi2Zp = (float)P.i2Zp; P.fi2Zs = (float)P.i2Zs;
    P.fcp = (float)P.cp; P.fcs = (float)P.cs;
    const double s9[9] = {0.0, 0.0, 0.0, h.sxx0, h.syy0, h.szz0, h.sxy0, h.sxz0, h.syz0};
    for (int s = 0; s < 2; ++s) {
        for (int c = 0; c < 9; ++c) P.bg[11 * s + c] = s9[c];
        P.bg[11 * s + 9] = 0.0;
        P.bg[11 * s + 10] = 0.0;
    }
    const double mbg = (h.sxx0 + h.syy0 + h.szz0) / 3.0;
    const double pbg[7] = {mbg, h.sxx0 - mbg, h.syy0 - mbg, h.szz0 - mbg, h.sxy0, h.sxz0, h.syz0};
    for (int k = 0; k < 7; ++k) P.pbg[k] = (float)pbg[k];
    const double pf[6] = {h.p_mu, h.p_b, h.p_h, h.p_h + h.G + h.p_mu * h.p_beta * P.K, h.p_eta, P.K * h.p_beta};
    for (int k = 0; k < 6; ++k) P.pf[k] = (float)pf[k];
    P.plast = h.use_plasticity;
    P.p_mu = h.p_mu; P.p_beta = h.p_beta; P.p_eta = h.p_eta; P.p_b = h.p_b; P.p_h = h.p_h;
    for (int k = 0; k < 4; ++k) P.bc[k] = h.bc[k];
    P.fault_kind = h.fault_kind;
    P.r0 = h.r0; P.nsig = h.nsig; P.rs_a = h.rs_a; P.rs_b = h.rs_b; P.rs_V0 = h.rs_V0; P.rs_f0 = h.rs_f0;
    P.rs_L = h.rs_L; P.rs_fw = h.rs_fw; P.rs_Vw = h.rs_Vw; P.rs_fexp = h.rs_fexp;
    P.x0 = h.x0; P.rs_thr = h.rs_rupture_thr; P.ramp_x = h.ramp_x; P.ramp_w = h.ramp_w; P.ramp_a = h.ramp_a;
    P.ramp_Vw = h.ramp_Vw;
    P.state_law = h.state_law;
    P.fs = h.fs; P.fd = h.fd; P.dw = h.dw; P.sm = h.sm_value; P.sw_thr = h.sw_rupture_thr;
    P.npatch = h.num_patches;
    for (int k = 0; k < TMK_MAXPATCH; ++k) { P.xc[k] = h.xc[k]; P.xw[k] = h.xw[k]; P.taup[k] = h.tau_p[k]; }
    P.fric_c0 = h.fric_center[0]; P.fric_w0 = h.fric_width[0]; P.outside_f = h.outside_f;
    P.face_lo = h.face_lo; P.nface = h.nface;
    P.pred = h.normal_pred; P.fourth = h.fourth_order; P.prim_lim = h.prim_limiting; P.char_lim = h.char_limiting;
    P.dx = h.dx;
    P.dlo[0] = h.domain.lo[0]; P.dlo[1] = h.domain.lo[1];
    P.dhi[0] = h.domain.hi[0]; P.dhi[1] = h.domain.hi[1];
}

// Streams of destroyed contexts, idle and synchronized, for the next contexts: creating and
// destroying a stream per scenario took milliseconds of host time in a serving loop.  Kept
// until the process ends (the list itself is never destroyed, so no static destructor runs
// after the HIP runtime has gone); every stream is non-blocking at the default priority.
static std::mutex g_stream_mu;
static std::vector<hipStream_t>& g_stream_pool = *new std::vector<hipStream_t>();

static hipError_t stream_take(hipStream_t* s)
{
    {
        std::lock_guard<std::mutex> lk(g_stream_mu);
        if (!g_stream_pool.empty()) {
            *s = g_stream_pool.back();
            g_stream_pool.pop_back();
            return hipSuccess;
        }
    }
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

static void stream_give(hipStream_t s)
{
    if (hipStreamSynchronize(s) != hipSuccess) {
        (void)hipStreamDestroy(s);
        return;
    }
    std::lock_guard<std::mutex> lk(g_stream_mu);
    if (g_stream_pool.size() < 64) g_stream_pool.push_back(s);
    else (void)hipStreamDestroy(s);
}

extern "C" int tmk_ctx_create(tmk_ctx** out, const tmk_params* p, int precision)
{
    if (!out || !p) return TMK_E_ARG;
    *out = nullptr;
    if (precision != TMK_PREC_FP64 && precision != TMK_PREC_FP32_BULK) return TMK_E_ARG;
    if (p->normal_pred < 0 || p->normal_pred > 2) return TMK_E_PRED_ORDER;
    if (p->normal_pred != 0 && p->char_limiting && p->prim_limiting) return TMK_E_PRED_ORDER;
    for (int k = 0; k < 4; ++k)
        if (k != 2 && p->bc[k] != 0 && p->bc[k] != 1) return TMK_E_BAD_BC;
    if (p->fault_kind != TMK_FAULT_RS && p->fault_kind != TMK_FAULT_SW) return TMK_E_ARG;
    if (p->num_patches < 0 || p->num_patches > TMK_MAXPATCH) return TMK_E_ARG;
    if (!(p->dx > 0.0) || p->domain.hi[0] - p->domain.lo[0] < 3 || p->domain.hi[1] - p->domain.lo[1] < 3)
        return TMK_E_ARG;
    tmk_ctx* c = new (std::nothrow) tmk_ctx;
    if (!c) return TMK_E_NOMEM;
    c->hp = *p;
    c->precision = precision;
    default_tile(precision, c->bx, c->by);
    fill_dev_params(*p, c->P);
    c->antiplane = !(p->sxz0 == 0.0 && p->syz0 == 0.0);
    if (stream_take(&c->stream) != hipSuccess) {
        c->stream = nullptr;
        delete c;
        return TMK_E_HIP;
    }
    // the allocation-free drop-in's scratch on this device; on failure only that entry point
    // is unavailable (it returns TMK_E_HIP)
    if (scratch_device_ptr(&c->noalloc_scratch) != TMK_OK) c->noalloc_scratch = nullptr;
    const double* fa[8] = {p->f_a, p->f_b, p->f_L, p->f_fs, p->f_fd, p->f_dc, p->f_syy, p->f_tau};
    const double** fd[8] = {&c->P.f_a, &c->P.f_b, &c->P.f_L, &c->P.f_fs, &c->P.f_fd, &c->P.f_dc, &c->P.f_syy,
                            &c->P.f_tau};
    // the per-face parameter arrays in one allocation, filled by one copy
    int nfa = 0;
    for (int k = 0; k < 8; ++k) {
        *fd[k] = nullptr;
        if (fa[k]) ++nfa;
    }
    if (nfa > 0) {
        if (p->nface <= 0) { tmk_ctx_destroy(c); return TMK_E_ARG; }
        const size_t n = (size_t)p->nface;
        std::vector<double> host(n * nfa);
        double* d = nullptr;
        if (dev_alloc(&d, sizeof(double) * host.size()) != hipSuccess) { tmk_ctx_destroy(c); return TMK_E_NOMEM; }
        c->face_dev.push_back(d);
        int q = 0;
        for (int k = 0; k < 8; ++k) {
            if (!fa[k]) continue;
            std::copy(fa[k], fa[k] + n, host.begin() + q * n);
            *fd[k] = d + q * n;
            ++q;
        }
        if (copy_on(c->stream, d, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice) != hipSuccess) {
            tmk_ctx_destroy(c);
            return TMK_E_HIP;
        }
    }
    c->hp.f_a = c->hp.f_b = c->hp.f_L = c->hp.f_fs = c->hp.f_fd = c->hp.f_dc = c->hp.f_syy = c->hp.f_tau = nullptr;
    *out = c;
    return TMK_OK;
}

extern "C" int tmk_ctx_destroy(tmk_ctx* c)
{
    if (!c) return TMK_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (double* d : c->face_dev) dev_free(d);  // the stream is idle and the levels are gone
    for (auto& b : c->bufs) b.release();
    if (c->stream) stream_give(c->stream);
    delete c;
    return TMK_OK;
}

static bool tile_ok(int bx, int by)
{
    return (bx == 16 && by == 16) || (bx == 32 && (by == 8 || by == 12 || by == 16 || by == 32)) ||
           (bx == 64 && (by == 4 || by == 8 || by == 16));
}

// Default tile per precision, from timing every shape on the 4096 x 2048 rate-and-state level on
// MI355X.  FP32 bulk: 32x16, one thread per cell.  FP64: 32x8, one thread per (cell, side), the
// fastest shape that keeps every value in registers; FP64 32x16 (one thread per cell) times
// 2-6 % faster on single levels but spills 52 B per thread at its 128-VGPR bound, which adds
// 0.5 GB of HBM writes per launch, and it is 5 % slower on levels cut into small boxes.
static void default_tile(int precision, int& bx, int& by)
{
    bx = 32;
    by = precision == TMK_PREC_FP64 ? TMK_DEFAULT_BY_F64 : 16;
}

extern "C" int tmk_set_march(tmk_ctx* c, int rows)
{
    if (!c || rows < 0 || rows > 65536) return TMK_E_ARG;
    c->march = rows;
    return TMK_OK;
}

extern "C" int tmk_set_tile(tmk_ctx* c, int bx, int by)
{
    if (!c) return TMK_E_ARG;
    if (bx == 0 && by == 0) default_tile(c->precision, bx, by);
    if (!tile_ok(bx, by)) return TMK_E_ARG;
    c->bx = bx;
    c->by = by;
    return TMK_OK;
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------
// predictor kind a scenario's numerics allow the kernel to be specialised for
static int pred_kind(const DevParams& P)
{
    return (P.pred == 2 && P.prim_lim && !P.char_lim) ? PK_PPM_PRIM : PK_RUNTIME;
}

template <typename T, int NS, int BX, int BY, int PK>
static int launch_tiles_t(hipStream_t st, const KArgs<T>& a)
{
    const size_t shm = interior_lds_elems<NS, BX, BY>() * sizeof(T);
    if (a.ntiles == 0) return TMK_OK;
    const void* k = (const void*)k_godunov_interior<T, NS, BX, BY, PK>;
    if (shm > 64 * 1024) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    hipLaunchKernelGGL((k_godunov_interior<T, NS, BX, BY, PK>), dim3(a.ntiles), dim3(interior_threads<T, NS, BX, BY>()),
                       shm, st, a);
    return hipGetLastError() == hipSuccess ? TMK_OK : TMK_E_HIP;
}

template <typename T, int NS, int PK>
static int launch_tiles_ns(hipStream_t st, int bx, int by, const KArgs<T>& a)
{
    if (bx == 32 && by == 16) return launch_tiles_t<T, NS, 32, 16, PK>(st, a);
    if (bx == 16 && by == 16) return launch_tiles_t<T, NS, 16, 16, PK>(st, a);
    if (bx == 32 && by == 8) return launch_tiles_t<T, NS, 32, 8, PK>(st, a);
    if (bx == 32 && by == 12) return launch_tiles_t<T, NS, 32, 12, PK>(st, a);
    if (bx == 32 && by == 32) return launch_tiles_t<T, NS, 32, 32, PK>(st, a);
    if (bx == 64 && by == 8) return launch_tiles_t<T, NS, 64, 8, PK>(st, a);
    if (bx == 64 && by == 4) return launch_tiles_t<T, NS, 64, 4, PK>(st, a);
    if (bx == 64 && by == 16) return launch_tiles_t<T, NS, 64, 16, PK>(st, a);
    return TMK_E_ARG;
}

// the boundary pieces a.tiles[0 .. a.ntiles) in one launch of k_godunov_bnd
template <typename T, int NS, int PK>
static int launch_bnd_t(hipStream_t st, const KArgs<T>& a)
{
    const size_t shm = bnd_lds_elems<T, NS>() * sizeof(T);
    if (a.ntiles == 0) return TMK_OK;
    const void* k = (const void*)k_godunov_bnd<T, NS, PK>;
    if (shm > 64 * 1024) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    hipLaunchKernelGGL((k_godunov_bnd<T, NS, PK>), dim3(a.ntiles), dim3(BND_BX * BND_BY), shm, st, a);
    return hipGetLastError() == hipSuccess ? TMK_OK : TMK_E_HIP;
}

template <typename T>
static int launch_bnd(hipStream_t st, int ns, int pk, const KArgs<T>& a)
{
    if (pk == PK_PPM_PRIM) return ns == 3 ? launch_bnd_t<T, 3, PK_PPM_PRIM>(st, a) : launch_bnd_t<T, 2, PK_PPM_PRIM>(st, a);
    return ns == 3 ? launch_bnd_t<T, 3, PK_RUNTIME>(st, a) : launch_bnd_t<T, 2, PK_RUNTIME>(st, a);
}

// one launch of k_godunov_interior over a.ntiles interior tiles of bx x by
template <typename T>
static int launch_tiles(hipStream_t st, int ns, int pk, int bx, int by, const KArgs<T>& a)
{
    if (pk == PK_PPM_PRIM)
        return ns == 3 ? launch_tiles_ns<T, 3, PK_PPM_PRIM>(st, bx, by, a) : launch_tiles_ns<T, 2, PK_PPM_PRIM>(st, bx, by, a);
    return ns == 3 ? launch_tiles_ns<T, 3, PK_RUNTIME>(st, bx, by, a) : launch_tiles_ns<T, 2, PK_RUNTIME>(st, bx, by, a);
}

// per-kernel timing of one step (profiling only): events around the interior launch on the
// main stream and around the boundary launch on the side stream
// Timing marks around the update launches of a step: HIP events (eager steps), or, inside a
// captured graph, a one-lane kernel that writes the device's constant-rate wall clock into
// stamps[0..3] (interior start, interior end, boundary start, boundary end).
struct StepTimers {
    hipEvent_t ia = nullptr, ib = nullptr, ba = nullptr, bb = nullptr;
    long long* stamps = nullptr;
};

static __global__ void k_stamp(long long* at)
{
    if (threadIdx.x == 0) *at = wall_clock64();
}

// mark k (0 interior start, 1 interior end, 2 boundary start, 3 boundary end) on stream s
static int timer_mark(const StepTimers* tm, hipStream_t s, int k)
{
    if (!tm) return TMK_OK;
    if (tm->stamps) {
        hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, s, tm->stamps + k);
        return hipGetLastError() == hipSuccess ? TMK_OK : TMK_E_HIP;
    }
    const hipEvent_t e = k == 0 ? tm->ia : k == 1 ? tm->ib : k == 2 ? tm->ba : tm->bb;
    return hipEventRecord(e, s) == hipSuccess ? TMK_OK : TMK_E_HIP;
}

// the fused update of one step over a tile list [interior | boundary]: the boundary launch runs
// on `side` (forked from and joined back into st by events; captured graphs keep the fork), the
// interior launch on st
// the interior strips of the row-marching partition (grid: two one-wavefront workgroups per strip);
// plast: some scenario of the launch has plasticity (the variant that reads and writes lam, gam)
template <typename T>
static int launch_march(hipStream_t st, const KArgs<T>& a, bool plast)
{
    if (plast && a.plastic)
        hipLaunchKernelGGL((k_godunov_march<T, PK_PPM_PRIM, true>), dim3(2 * a.ntiles), dim3(64), 0, st, a);
    else
        hipLaunchKernelGGL((k_godunov_march<T, PK_PPM_PRIM, false>), dim3(2 * a.ntiles), dim3(64), 0, st, a);
    return hipGetLastError() == hipSuccess ? TMK_OK : TMK_E_HIP;
}

template <typename T>
static int launch_update(hipStream_t st, hipStream_t side, hipEvent_t fork, hipEvent_t join, int ns, int pk, int bx,
                         int by, KArgs<T> a, int nint, const StepTimers* tm, int march = 0, bool plast = true)
{
    const int nbnd = a.ntiles - nint;
    const Tile* tiles = a.tiles;
    int e = TMK_OK;
    const bool split = nbnd > 0 && nint > 0 && side;
    if (split) {
        HIPCHK(hipEventRecord(fork, st));
        HIPCHK(hipStreamWaitEvent(side, fork, 0));
    }
    if (nbnd > 0) {
        hipStream_t sb = split ? side : st;
        KArgs<T> b = a;
        b.tiles = tiles + nint;
        b.ntiles = nbnd;
        if ((e = timer_mark(tm, sb, 2))) return e;
        if ((e = launch_bnd<T>(sb, ns, pk, b))) return e;
        if ((e = timer_mark(tm, sb, 3))) return e;
    }
    if (nint > 0) {
        KArgs<T> b = a;
        b.ntiles = nint;
        if ((e = timer_mark(tm, st, 0))) return e;
        if (march > 0) {
            if ((e = launch_march<T>(st, b, plast))) return e;
        } else if ((e = launch_tiles<T>(st, ns, pk, bx, by, b))) {
            return e;
        }
        if ((e = timer_mark(tm, st, 1))) return e;
    }
    if (split) {
        HIPCHK(hipEventRecord(join, side));
        HIPCHK(hipStreamWaitEvent(st, join, 0));
    }
    return TMK_OK;
}

// one level step of every member of a group: fused update, ghost exchange of the
// new state, clock advance
template <typename T>
static int group_enqueue_step_t(Group& G, hipStream_t st, const StepTimers* tm = nullptr)
{
    KArgs<T> a;
    a.scen = G.scen_dev;
    a.patches = (const Patch<T>*)G.patches_dev[G.g];
    a.tiles = G.tiles_dev;
    a.ntiles = G.ntiles;
    a.plastic = G.plastic;
    int e = launch_update<T>(st, G.side, G.ev_fork, G.ev_join, G.ns, G.pk, G.bx, G.by, a, G.nint, tm, G.march,
                             G.any_plast);
    if (e) return e;
    const Patch<T>* pn = (const Patch<T>*)G.patches_dev[G.g];  // its out planes hold the new state
    if (G.nrects > 0) {
        hipLaunchKernelGGL(k_exchange<T>, dim3(G.nrects), dim3(128), 0, st, pn, G.rects_dev, G.nrects,
                           stencil_mask(G.ns));
        if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
    }
    if (G.nbrects > 0) {
        hipLaunchKernelGGL(k_bexchange<T>, dim3(G.nbrects), dim3(64), 0, st, pn, G.brects_dev, G.nbrects);
        if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
    }
    hipLaunchKernelGGL(k_advance_time, dim3((G.nscen + 63) / 64), dim3(64), 0, st, G.scen_dev, G.nscen);
    if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
    if (G.stations) {
        hipLaunchKernelGGL(k_station_record<T>, dim3(1), dim3(256), 0, st, pn, G.scen_dev, *G.stations);
        if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
    }
    G.g ^= 1;
    return TMK_OK;
}

static int group_enqueue_step(Group& G, hipStream_t st, const StepTimers* tm = nullptr)
{
    return G.tsize == sizeof(double) ? group_enqueue_step_t<double>(G, st, tm) : group_enqueue_step_t<float>(G, st, tm);
}

// exchange the current state (before the first step after an upload)
template <typename T>
static int group_exchange_now_t(Group& G, hipStream_t st)
{
    // patches_dev[g ^ 1].out are the planes holding the current state
    const Patch<T>* pn = (const Patch<T>*)G.patches_dev[G.g ^ 1];
    if (G.nrects > 0) {
        hipLaunchKernelGGL(k_exchange<T>, dim3(G.nrects), dim3(128), 0, st, pn, G.rects_dev, G.nrects,
                           stencil_mask(G.ns));
        if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
    }
    if (G.nbrects > 0) {
        hipLaunchKernelGGL(k_bexchange<T>, dim3(G.nbrects), dim3(64), 0, st, pn, G.brects_dev, G.nbrects);
        if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
    }
    return TMK_OK;
}

// n steps; a two-step graph (parities 0 -> 1 -> 0) is captured once per group
// layout and dt and replayed
static int group_run(Group& G, hipStream_t st, int nsteps)
{
    int e;
    if (nsteps <= 0) return TMK_OK;
    if (G.g != 0) {
        if ((e = group_enqueue_step(G, st))) return e;
        --nsteps;
    }
    if (nsteps >= 2 && !G.gexec) {
        hipGraph_t gr;
        const int g0 = G.g;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        e = group_enqueue_step(G, st);
        if (!e) e = group_enqueue_step(G, st);
        hipError_t ce = hipStreamEndCapture(st, &gr);
        G.g = g0;  // nothing ran; two captured steps return to this parity, a failed capture must too
        if (e) return e;
        if (ce != hipSuccess) return TMK_E_HIP;
        hipError_t ie = hipGraphInstantiate(&G.gexec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ie != hipSuccess) return TMK_E_HIP;
    }
    for (int k = 0; k + 1 < nsteps; k += 2) HIPCHK(hipGraphLaunch(G.gexec, st));
    if (nsteps % 2)
        if ((e = group_enqueue_step(G, st))) return e;
    return TMK_OK;
}

// Build a group from member levels.  Member m's buffer parity at group parity g
// is phase[m] ^ g.
template <typename T>
static int group_build_t(Group& G, const std::vector<tmk_level*>& mem, const std::vector<int>& phase, int plastic)
{
    G.reset();
    if (mem.empty()) return TMK_OK;
    tmk_level* L0 = mem[0];
    G.tsize = sizeof(T);
    G.ns = L0->ns;
    G.pk = PK_PPM_PRIM;
    for (tmk_level* L : mem)
        if (pred_kind(L->ctx->P) != PK_PPM_PRIM) G.pk = PK_RUNTIME;
    G.bx = L0->bx;
    G.by = L0->by;
    G.march = (L0->march > 0 && G.ns == 2 && G.pk == PK_PPM_PRIM) ? L0->march : 0;
    G.xcd = L0->xcd;
    G.nb = L0->ctx->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
    G.plastic = plastic;
    G.any_plast = false;
    for (tmk_level* L : mem) G.any_plast = G.any_plast || L->ctx->P.plast;
    G.g = 0;
    std::vector<Scen> sc(mem.size());
    std::vector<Patch<T>> pts[2];
    std::vector<Tile> tiles;
    std::vector<CopyRect> rects, brects;
    int poff = 0;
    for (size_t m = 0; m < mem.size(); ++m) {
        tmk_level* L = mem[m];
        memset(&sc[m], 0, sizeof(Scen));
        sc[m].P = L->ctx->P;
        scen_set_clock(sc[m], L->time, L->dt);
        const int np = L->npatch;
        for (int g = 0; g < 2; ++g) {
            const Patch<T>* src = (const Patch<T>*)L->hpatch[phase[m] ^ g].data();
            for (int b = 0; b < np; ++b) {
                Patch<T> p = src[b];
                p.scen = (int)m;
                pts[g].push_back(p);
            }
        }
        if (G.march) {
            for (Tile t : L->urects) { t.patch += poff; tiles.push_back(t); }
        } else {
            for (Tile t : L->tiles) { t.patch += poff; tiles.push_back(t); }
        }
        for (CopyRect r : L->rects) { r.src += poff; r.dst += poff; rects.push_back(r); }
        for (CopyRect r : L->brects) { r.src += poff; r.dst += poff; brects.push_back(r); }
        poff += np;
    }
    // tile -> scenario through the patch index (patches of member m follow those of m - 1)
    std::vector<int> patch_scen;
    for (size_t m = 0; m < mem.size(); ++m)
        for (int b = 0; b < mem[m]->npatch; ++b) patch_scen.push_back((int)m);
    const int bx = G.bx, by = G.by;
    if (G.march) {
        // the rectangles become interior strips, then boundary pieces (fault row first)
        std::vector<Tile> strips, bf, bo;
        // strips of levels with flux registers hold at most 63 rows (the kernel's face masks)
        bool fluxes = false;
        for (const Patch<T>& p : pts[0]) fluxes = fluxes || p.F[0] || p.F[1];
        const int rows = fluxes ? std::min(G.march, 63) : G.march;
        for (const Tile& r : tiles) march_rect(r, sc[patch_scen[r.patch]].P, rows, strips, bf, bo);
        if (G.xcd) xcd_order(strips);
        G.nint = (int)strips.size();
        tiles.swap(strips);
        tiles.insert(tiles.end(), bf.begin(), bf.end());
        tiles.insert(tiles.end(), bo.begin(), bo.end());
    } else {
        G.nint = partition_tiles(
            tiles, [&](const Tile& t) { return tile_interior_host(t, sc[patch_scen[t.patch]].P, bx, by); },
            [&](const Tile& t) -> const DevParams& { return sc[patch_scen[t.patch]].P; }, G.xcd != 0);
    }
    G.cells_int = G.cells_bnd = 0;
    for (size_t k = 0; k < tiles.size(); ++k)
        ((int)k < G.nint ? G.cells_int : G.cells_bnd) +=
            (long long)(tiles[k].x1 - tiles[k].x0 + 1) * (tiles[k].y1 - tiles[k].y0 + 1);
    if (!G.side) {
        HIPCHK(hipStreamCreateWithFlags(&G.side, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&G.ev_fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&G.ev_join, hipEventDisableTiming));
    }
    // the tables are rewritten in place: every launch that read them has completed
    // (the callers synchronise the group's stream first; the side stream joins it)
    auto put = [&G](DevBuf& buf, const void* src, size_t bytes, void** dst) -> int {
        int e = buf.ensure(std::max<size_t>(bytes, 1));
        if (e) return e;
        *dst = buf.p;
        if (copy_on(G.side, buf.p, src, bytes, hipMemcpyHostToDevice) != hipSuccess) return TMK_E_HIP;
        return TMK_OK;
    };
    int e;
    if ((e = put(G.mem_scen, sc.data(), sizeof(Scen) * sc.size(), (void**)&G.scen_dev))) return e;
    for (int g = 0; g < 2; ++g)
        if ((e = put(G.mem_pt[g], pts[g].data(), sizeof(Patch<T>) * pts[g].size(), &G.patches_dev[g]))) return e;
    if ((e = put(G.mem_tiles, tiles.data(), sizeof(Tile) * tiles.size(), (void**)&G.tiles_dev))) return e;
    if ((e = put(G.mem_rects, rects.data(), sizeof(CopyRect) * rects.size(), (void**)&G.rects_dev))) return e;
    if ((e = put(G.mem_brects, brects.data(), sizeof(CopyRect) * brects.size(), (void**)&G.brects_dev))) return e;
    G.nscen = (int)sc.size();
    G.ntiles = (int)tiles.size();
    G.nrects = (int)rects.size();
    G.nbrects = (int)brects.size();
    return TMK_OK;
}

static int group_build(Group& G, const std::vector<tmk_level*>& mem, const std::vector<int>& phase, int plastic)
{
    if (mem.empty()) { G.reset(); return TMK_OK; }
    return mem[0]->tsize == sizeof(double) ? group_build_t<double>(G, mem, phase, plastic)
                                           : group_build_t<float>(G, mem, phase, plastic);
}

// the clock and step constants (Scen fields time .. fmdtdx) of member m
static int group_set_clock(Group& G, hipStream_t st, int m, double time, double dt, const DevParams& P)
{
    Scen s;
    s.P = P;
    scen_set_clock(s, time, dt);
    const size_t off = offsetof(Scen, time), len = offsetof(Scen, err) - off;
    // pageable source: the call returns once s has been staged, so s may go out of scope
    HIPCHK(hipMemcpyAsync((char*)(G.scen_dev + m) + off, (const char*)&s + off, len, hipMemcpyHostToDevice, st));
    return TMK_OK;
}

static int group_error(Group& G, hipStream_t st, int m, int* code)
{
    int e = 0;
    HIPCHK(copy_on(st, &e, (char*)(G.scen_dev + m) + offsetof(Scen, err), sizeof(int), hipMemcpyDeviceToHost));
    *code = e;
    return TMK_OK;
}

// ---------------------------------------------------------------------------
// per-patch drop-in (host FArrayBox data)
// ---------------------------------------------------------------------------

// The tiles of one box for the drop-in path, with no heap allocation: interior tiles and the
// cores of the others, then those tiles' bands in boundary pieces, fault-row pieces first -- the
// partition of partition_tiles (without its XCD ordering, which a single box does not need; the
// order of the tiles does not change any result).  Writes at most cap tiles to out (out may be NULL to
// count); returns the count, and the interior count in *nint.
static int patch_tiles(const tmk_box& v, const DevParams& P, int bx, int by, Tile* out, int cap, int* nint)
{
    const int tx = bx - 2, ty = by - 2;
    int n = 0;
    auto put = [&](const Tile& t) {
        if (out && n < cap) out[n] = t;
        ++n;
    };
    for (int pass = 0; pass < 3; ++pass) {
        for (int y0 = v.lo[1]; y0 <= v.hi[1]; y0 += ty)
            for (int x0 = v.lo[0]; x0 <= v.hi[0]; x0 += tx) {
                Tile t;
                t.patch = 0;
                t.x0 = x0;
                t.y0 = y0;
                t.x1 = std::min(x0 + tx - 1, v.hi[0]);
                t.y1 = std::min(y0 + ty - 1, v.hi[1]);
                const bool in = tile_interior_host(t, P, bx, by);
                if (pass == 0) {
                    if (in) {
                        put(t);
                    } else {
                        const Tile c = band_split(t, P, [](const Tile&) {});
                        if (!tile_empty(c)) put(c);
                    }
                    continue;
                }
                if (in) continue;
                (void)band_split(t, P, [&](const Tile& p) {
                    if ((p.y0 == P.dlo[1]) == (pass == 1)) put(p);
                });
            }
        if (pass == 0) *nint = n;
    }
    return n;
}

// FP64 state <-> the FP32 bulk's stored form U - bg (stores_delta), component blockIdx.y
static __global__ void k_to_f32(const double* __restrict__ s, float* __restrict__ d, size_t plane, StateVec bg)
{
    const size_t o = blockIdx.y * plane;
    const double b = bg.c[blockIdx.y];
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < plane; k += (size_t)gridDim.x * blockDim.x)
        d[o + k] = (float)(s[o + k] - b);
}
static __global__ void k_to_f64(const float* __restrict__ s, double* __restrict__ d, size_t plane, StateVec bg)
{
    const size_t o = blockIdx.y * plane;
    const double b = bg.c[blockIdx.y];
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < plane; k += (size_t)gridDim.x * blockDim.x)
        d[o + k] = (double)s[o + k] + b;
}

// the background per component (the FP32 bulk stores U minus it)
static StateVec background(const DevParams& P)
{
    StateVec v;
    for (int k = 0; k < TMK_NCOMP; ++k) v.c[k] = P.bg[k];
    return v;
}
// GETFLUXF (LinElastPhysicsF.ChF:303-350) of the background in direction d, per component: the
// FP32 bulk's flux outputs are fluxes of U - bg, and the drop-in adds this to return the flux
// of U (the same slot forms as the kernels' face_flux_side and store_flux_side)
static void bg_flux(const DevParams& P, int d, double out[TMK_NCOMP])
{
    for (int sd = 0; sd < 2; ++sd) {
        const int o = 11 * sd;
        const double r = (d == 1 && sd == 1) ? -1.0 : 1.0;
        double v[3], sv[3];
        for (int k = 0; k < 3; ++k) {
            v[k] = P.bg[slot_comp(d, k, 0) + o];
            sv[k] = P.bg[slot_comp(d, k, 1) + o];
        }
        const double f0 = -(r * sv[0]) / P.rho, f2 = -(r * sv[1]) / P.rho, f5 = -(r * sv[2]) / P.rho;
        const double f1 = -(r * v[0] * P.lam2G), f4 = -(r * v[0] * P.lam), f3 = -(r * v[1] * P.G);
        const double f6 = -(r * v[2] * P.G);
        double* q = out + o;
        if (d == 0) {
            q[0] = f0; q[1] = f2; q[2] = f5; q[3] = f1; q[4] = f4; q[5] = f4; q[6] = f3; q[7] = f6; q[8] = 0.0;
        } else {
            q[1] = f0; q[0] = f2; q[2] = f5; q[4] = f1; q[5] = f4; q[3] = f4; q[6] = f3; q[8] = f6; q[7] = 0.0;
        }
        q[9] = q[10] = 0.0;
    }
}
static void add_bg_flux(const DevParams& P, int d, double* F, size_t n)
{
    double fb[TMK_NCOMP];
    bg_flux(P, d, fb);
    const size_t plane = n / TMK_NCOMP;
    for (int c = 0; c < TMK_NCOMP; ++c)
        if (fb[c] != 0.0)
            for (size_t k = 0; k < plane; ++k) F[k + c * plane] += fb[c];
}

// host copies between FP64 state and the FP32 bulk's stored form
static void to_stored_f32(const double* U, float* d, size_t plane, const DevParams& P)
{
    for (int c = 0; c < TMK_NCOMP; ++c)
        for (size_t k = 0; k < plane; ++k) d[k + c * plane] = (float)(U[k + c * plane] - P.bg[c]);
}
static void from_stored_f32(const float* a, double* U, size_t plane, const DevParams& P)
{
    for (int c = 0; c < TMK_NCOMP; ++c)
        for (size_t k = 0; k < plane; ++k) U[k + c * plane] = (double)a[k + c * plane] + P.bg[c];
}

// Device buffers of one drop-in call.  stage: FP64 image of the state for the FP32 bulk,
// converted on the device (NULL: converted on the host).  meta: the patch descriptor, the
// tiles and the scenario block, at the offsets patch_meta gives.
template <typename T>
struct PatchWork {
    T* Uin = nullptr;
    T* Uout = nullptr;
    double* Bd = nullptr;
    double* Fd = nullptr;
    double* stage = nullptr;
    char* meta = nullptr;
};

template <typename T>
static void patch_meta(int ntiles, size_t* off_t, size_t* off_s, size_t* bytes)
{
    *off_t = ((sizeof(Patch<T>) + 255) / 256) * 256;
    *off_s = *off_t + ((sizeof(Tile) * (size_t)ntiles + 255) / 256) * 256;
    *bytes = *off_s + sizeof(Scen);
}

// Sizes of one drop-in call, in elements
struct PatchSizes {
    size_t unx, plane, nU, bnx, nB, nF0, nF1;
};
static PatchSizes patch_sizes(const tmk_box& valid, const tmk_box& ub, const double* F0, const double* F1,
                              const double* B, const tmk_box& bb, int nb)
{
    PatchSizes z;
    z.unx = (size_t)(ub.hi[0] - ub.lo[0] + 1);
    z.plane = z.unx * (size_t)(ub.hi[1] - ub.lo[1] + 1);
    z.nU = z.plane * TMK_NCOMP;
    z.bnx = B ? (size_t)(bb.hi[0] - bb.lo[0] + 1) : 0;
    z.nB = z.bnx * (size_t)nb;
    z.nF0 = F0 ? (size_t)(valid.hi[0] - valid.lo[0] + 2) * (valid.hi[1] - valid.lo[1] + 1) * TMK_NCOMP : 0;
    z.nF1 = F1 ? (size_t)(valid.hi[0] - valid.lo[0] + 1) * (valid.hi[1] - valid.lo[1] + 2) * TMK_NCOMP : 0;
    return z;
}

// One drop-in call on the buffers of w.  tiles: the box's tiles (patch_tiles), nint interior.
template <typename T>
static int patch_update_t(tmk_ctx* c, tmk_box valid, double* U, tmk_box ub, double* F0, double* F1, double* B,
                          tmk_box bb, int nb, double time, double dt, bool plastic_only, const PatchWork<T>& w,
                          const Tile* tiles, int ntiles, int nint)
{
    const PatchSizes z = patch_sizes(valid, ub, F0, F1, B, bb, nb);
    const size_t unx = z.unx, plane = z.plane, nU = z.nU, bnx = z.bnx, nB = z.nB, nF0 = z.nF0, nF1 = z.nF1;
    if (plane * sizeof(T) > kPlaneMaxBytes) return TMK_E_NOMEM;  // 32-bit plane offsets (gld/gst)
    hipStream_t st = c->stream;
    T* Uin = w.Uin;
    T* Uout = w.Uout;
    double* Bd = w.Bd;
    double* Fd = w.Fd;
    const dim3 cvt_grid((unsigned)std::min<size_t>((plane + 255) / 256, 1024), TMK_NCOMP);
    std::vector<float> tmp;
    if (sizeof(T) == sizeof(double)) {
        HIPCHK(hipMemcpyAsync(Uin, U, nU * sizeof(double), hipMemcpyHostToDevice, st));
    } else if (w.stage) {
        HIPCHK(hipMemcpyAsync(w.stage, U, nU * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_to_f32, cvt_grid, dim3(256), 0, st, (const double*)w.stage, (float*)Uin, plane,
                           background(c->P));
        HIPCHK(hipGetLastError());
    } else {
        tmp.resize(nU);
        to_stored_f32(U, tmp.data(), plane, c->P);
        HIPCHK(hipMemcpyAsync(Uin, tmp.data(), nU * sizeof(float), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(Uout, Uin, nU * sizeof(T), hipMemcpyDeviceToDevice, st));
    if (plastic_only) {
        const int vnx = valid.hi[0] - valid.lo[0] + 1, vny = valid.hi[1] - valid.lo[1] + 1;
        const int n = vnx * vny;
        hipLaunchKernelGGL(k_plastic<T>, dim3((n + 255) / 256), dim3(256), 0, st, c->P, Uout, ub.lo[0], ub.lo[1],
                           (int)unx, plane, valid.lo[0], valid.lo[1], vnx, vny, dt);
        HIPCHK(hipGetLastError());
    } else {
        if (B) {
            HIPCHK(hipMemcpyAsync(Bd, B, nB * sizeof(double), hipMemcpyHostToDevice, st));
            HIPCHK(hipMemcpyAsync(Bd + nB, Bd, nB * sizeof(double), hipMemcpyDeviceToDevice, st));
        }
        Patch<T> pt;
        memset(&pt, 0, sizeof pt);
        for (int k = 0; k < TMK_NCOMP; ++k) { pt.in[k] = Uin + k * plane; pt.out[k] = Uout + k * plane; }
        pt.ulo0 = ub.lo[0]; pt.ulo1 = ub.lo[1]; pt.unx = (int)unx;
        pt.pbytes = (unsigned)(plane * sizeof(T));
        pt.bin = B ? Bd : nullptr;
        pt.bout = B ? Bd + nB : nullptr;
        pt.blo0 = bb.lo[0]; pt.bnx = (int)bnx; pt.bnb = nb;
        pt.F[0] = F0 ? Fd : nullptr;
        pt.F[1] = F1 ? Fd + nF0 : nullptr;
        pt.vlo0 = valid.lo[0]; pt.vlo1 = valid.lo[1]; pt.vhi0 = valid.hi[0]; pt.vhi1 = valid.hi[1];
        pt.scen = 0;
        size_t off_t, off_s, mbytes;
        patch_meta<T>(ntiles, &off_t, &off_s, &mbytes);
        char* meta = w.meta;
        Scen sc;
        memset(&sc, 0, sizeof sc);
        sc.P = c->P;
        scen_set_clock(sc, time, dt);
        HIPCHK(hipMemcpyAsync(meta, &pt, sizeof pt, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(meta + off_t, tiles, sizeof(Tile) * ntiles, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(meta + off_s, &sc, sizeof sc, hipMemcpyHostToDevice, st));
        KArgs<T> a;
        a.scen = (Scen*)(meta + off_s);
        a.patches = (const Patch<T>*)meta;
        a.tiles = (const Tile*)(meta + off_t);
        a.ntiles = ntiles;
        a.plastic = 0;  // updateState does not include the plastic return
        // the anti-plane planes are skipped only when background and data are zero there
        int ns = 3;
        if (!c->antiplane) {
            bool zero = true;
            for (int s = 0; s < 2 && zero; ++s)
                for (int cc : {2, 7, 8}) {
                    const double* q = U + (size_t)(11 * s + cc) * plane;
                    for (size_t k = 0; k < plane; ++k)
                        if (q[k] != 0.0) { zero = false; break; }
                    if (!zero) break;
                }
            ns = zero ? 2 : 3;
        }
        int e;
        if ((e = launch_update<T>(st, nullptr, nullptr, nullptr, ns, pred_kind(c->P), c->bx, c->by, a, nint, nullptr)))
            return e;
        int herr = 0;
        HIPCHK(hipMemcpyAsync(&herr, meta + off_s + offsetof(Scen, err), sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (herr) return herr;
        if (B) HIPCHK(copy_on(st, B, Bd + nB, nB * sizeof(double), hipMemcpyDeviceToHost));
        if (F0) HIPCHK(copy_on(st, F0, Fd, nF0 * sizeof(double), hipMemcpyDeviceToHost));
        if (F1) HIPCHK(copy_on(st, F1, Fd + nF0, nF1 * sizeof(double), hipMemcpyDeviceToHost));
        if (stores_delta<T>()) {
            if (F0) add_bg_flux(c->P, 0, F0, nF0);
            if (F1) add_bg_flux(c->P, 1, F1, nF1);
        }
    }
    if (sizeof(T) == sizeof(double)) {
        HIPCHK(hipMemcpyAsync(U, Uout, nU * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    } else if (w.stage) {
        hipLaunchKernelGGL(k_to_f64, cvt_grid, dim3(256), 0, st, (const float*)Uout, w.stage, plane,
                           background(c->P));
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(U, w.stage, nU * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    } else {
        tmp.resize(nU);
        HIPCHK(hipMemcpyAsync(tmp.data(), Uout, nU * sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        from_stored_f32(tmp.data(), U, plane, c->P);
    }
    return TMK_OK;
}

// drop-in call on the context's grow-only buffers (allocated on first need)
template <typename T>
static int patch_update_bufs(tmk_ctx* c, tmk_box valid, double* U, tmk_box ub, double* F0, double* F1, double* B,
                             tmk_box bb, int nb, double time, double dt, bool plastic_only)
{
    const PatchSizes z = patch_sizes(valid, ub, F0, F1, B, bb, nb);
    if (z.plane * sizeof(T) > kPlaneMaxBytes) return TMK_E_NOMEM;
    int nint = 0;
    std::vector<Tile> tiles;
    if (!plastic_only) {
        const int n = patch_tiles(valid, c->P, c->bx, c->by, nullptr, 0, &nint);
        tiles.resize(n);
        patch_tiles(valid, c->P, c->bx, c->by, tiles.data(), n, &nint);
    }
    size_t off_t, off_s, mbytes;
    patch_meta<T>((int)tiles.size(), &off_t, &off_s, &mbytes);
    int e;
    if ((e = c->bufs[1].ensure(z.nU * sizeof(T)))) return e;
    if ((e = c->bufs[2].ensure(z.nU * sizeof(T)))) return e;
    if ((e = c->bufs[3].ensure(std::max<size_t>(2 * z.nB, 1) * sizeof(double)))) return e;
    if ((e = c->bufs[4].ensure(std::max<size_t>(z.nF0 + z.nF1, 1) * sizeof(double)))) return e;
    if ((e = c->bufs[5].ensure(mbytes))) return e;
    PatchWork<T> w;
    w.Uin = (T*)c->bufs[1].p;
    w.Uout = (T*)c->bufs[2].p;
    w.Bd = (double*)c->bufs[3].p;
    w.Fd = (double*)c->bufs[4].p;
    w.meta = (char*)c->bufs[5].p;
    return patch_update_t<T>(c, valid, U, ub, F0, F1, B, bb, nb, time, dt, plastic_only, w, tiles.data(),
                             (int)tiles.size(), nint);
}

// argument checks of tmk_patch_update_state (B is cleared when valid does not touch the fault)
static int patch_args(tmk_ctx* c, const tmk_box& valid, const double* U, const tmk_box& ub, int ncomp, double*& B,
                      const tmk_box& bb, int nbdry)
{
    if (!c || !U || ncomp != TMK_NCOMP) return TMK_E_ARG;
    if (box_empty(valid)) return TMK_E_ARG;
    const tmk_box dom = c->hp.domain;
    if (!box_contains(dom, valid)) return TMK_E_BOX;
    // the update reads W on grow(valid,3) & domain and writes the valid cells
    if (!box_contains(ub, box_and(box_grow(valid, 3), dom))) return TMK_E_BOX;
    const bool fault = valid.lo[1] == dom.lo[1];
    if (fault) {
        if (!B) return TMK_E_NO_BDRY;
        const int need = c->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
        if (nbdry != need) return TMK_E_ARG;
        const int lo = std::max(valid.lo[0] - 1, dom.lo[0]);
        const int hi = std::min(valid.hi[0] + 1, dom.hi[0]);
        if (bb.lo[0] > lo || bb.hi[0] < hi) return TMK_E_BOX;
    } else {
        B = nullptr;
    }
    return TMK_OK;
}

extern "C" int tmk_patch_update_state(tmk_ctx* c, tmk_box valid, double* U, tmk_box ub, int ncomp, double* F0,
                                      double* F1, double* B, tmk_box bb, int nbdry, double time, double dt,
                                      double* max_wave_speed)
{
    int e;
    if ((e = patch_args(c, valid, U, ub, ncomp, B, bb, nbdry))) return e;
    if (max_wave_speed) *max_wave_speed = std::max(c->hp.cs, c->hp.cp);
    if (c->precision == TMK_PREC_FP64)
        return patch_update_bufs<double>(c, valid, U, ub, F0, F1, B, bb, nbdry, time, dt, false);
    return patch_update_bufs<float>(c, valid, U, ub, F0, F1, B, bb, nbdry, time, dt, false);
}

extern "C" int tmk_patch_plastic_update(tmk_ctx* c, tmk_box valid, double* U, tmk_box ub, int ncomp, double dt)
{
    if (!c || !U || ncomp != TMK_NCOMP) return TMK_E_ARG;
    if (box_empty(valid) || !box_contains(ub, valid)) return TMK_E_BOX;
    if (c->precision == TMK_PREC_FP64)
        return patch_update_bufs<double>(c, valid, U, ub, nullptr, nullptr, nullptr, ub, 0, 0.0, dt, true);
    return patch_update_bufs<float>(c, valid, U, ub, nullptr, nullptr, nullptr, ub, 0, 0.0, dt, true);
}

// ---------------------------------------------------------------------------
// allocation-free drop-in
// ---------------------------------------------------------------------------
// Device scratch of tmk_patch_update_state_noalloc: the two state copies, the fault data (read
// and written copies), the fluxes, the FP64 staging of the FP32 bulk and the launch tables of
// one call.  One array per device for the life of the process; every call on a device shares
// it, so calls that overlap on one device race.  Host code never passes this name to a kernel
// or a copy: it is a host-side stand-in, and the device address comes from
// hipGetSymbolAddress (scratch_device_ptr).
static __device__ __align__(256) unsigned char scratch[TMK_NOALLOC_SCRATCH_BYTES];

// the most tiles one allocation-free call launches (they are listed in a stack array)
static constexpr int kNoAllocMaxTiles = 1024;

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];  // per device ordinal, resolved once

// device address of scratch on the current device (resolved once per device, at context set-up)
static int scratch_device_ptr(void** out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return TMK_E_HIP;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "tetemoko_hip: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return TMK_E_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TMK_OK;
}

static size_t align256(size_t b) { return ((b + 255) / 256) * 256; }

template <typename T>
static int patch_update_noalloc_t(tmk_ctx* c, tmk_box valid, double* U, tmk_box ub, double* F0, double* F1,
                                  double* B, tmk_box bb, int nb, double time, double dt)
{
    if (!c->noalloc_scratch) return TMK_E_HIP;
    const PatchSizes z = patch_sizes(valid, ub, F0, F1, B, bb, nb);
    int nint = 0;
    const int ntiles = patch_tiles(valid, c->P, c->bx, c->by, nullptr, 0, &nint);
    if (ntiles > kNoAllocMaxTiles) return TMK_E_SCRATCH;
    size_t off_t, off_s, mbytes;
    patch_meta<T>(ntiles, &off_t, &off_s, &mbytes);
    // layout: Uin, Uout, fault data (2 copies), fluxes, FP64 staging (FP32 bulk only), tables
    const size_t oUout = align256(z.nU * sizeof(T));
    const size_t oB = oUout + align256(z.nU * sizeof(T));
    const size_t oF = oB + align256(2 * z.nB * sizeof(double));
    const size_t oS = oF + align256((z.nF0 + z.nF1) * sizeof(double));
    const size_t oM = oS + (sizeof(T) == sizeof(double) ? 0 : align256(z.nU * sizeof(double)));
    if (oM + mbytes > (size_t)TMK_NOALLOC_SCRATCH_BYTES) return TMK_E_SCRATCH;
    Tile tiles[kNoAllocMaxTiles];
    patch_tiles(valid, c->P, c->bx, c->by, tiles, kNoAllocMaxTiles, &nint);
    char* base = (char*)c->noalloc_scratch;
    PatchWork<T> w;
    w.Uin = (T*)base;
    w.Uout = (T*)(base + oUout);
    w.Bd = (double*)(base + oB);
    w.Fd = (double*)(base + oF);
    w.stage = sizeof(T) == sizeof(double) ? nullptr : (double*)(base + oS);
    w.meta = base + oM;
    return patch_update_t<T>(c, valid, U, ub, F0, F1, B, bb, nb, time, dt, false, w, tiles, ntiles, nint);
}

extern "C" int tmk_patch_update_state_noalloc(tmk_ctx* c, tmk_box valid, double* U, tmk_box ub, int ncomp,
                                              double* F0, double* F1, double* B, tmk_box bb, int nbdry, double time,
                                              double dt, double* max_wave_speed)
{
    int e;
    if ((e = patch_args(c, valid, U, ub, ncomp, B, bb, nbdry))) return e;
    if (max_wave_speed) *max_wave_speed = std::max(c->hp.cs, c->hp.cp);
    if (c->precision == TMK_PREC_FP64)
        return patch_update_noalloc_t<double>(c, valid, U, ub, F0, F1, B, bb, nbdry, time, dt);
    return patch_update_noalloc_t<float>(c, valid, U, ub, F0, F1, B, bb, nbdry, time, dt);
}

// ---------------------------------------------------------------------------
// device-resident level
// ---------------------------------------------------------------------------
template <typename T>
static void level_fill_patches(tmk_level* L)
{
    const int nbox = (int)L->boxes.size(), nisl = (int)L->isl.size();
    L->npatch = nbox + nisl;
    for (int cur = 0; cur < 2; ++cur) {
        L->hpatch[cur].assign(sizeof(Patch<T>) * L->npatch, 0);
        Patch<T>* pts = (Patch<T>*)L->hpatch[cur].data();
        for (int q = 0; q < L->npatch; ++q) {
            const int k = q < nbox ? L->box_isl[q] : q - nbox;
            const Island& I = L->isl[k];
            const tmk_box v = q < nbox ? L->boxes[q] : I.bbox;
            const bool onf = q < nbox ? L->has_b[q] != 0 : I.has_b != 0;
            const size_t plane = I.plane();
            Patch<T>& p = pts[q];
            T* A = (T*)I.u[cur];
            T* Bn = (T*)I.u[1 - cur];
            T* W = (T*)I.w;
            for (int c = 0; c < TMK_NCOMP; ++c) {
                if (is_stencil(c, L->ns)) { p.in[c] = A + c * plane; p.out[c] = Bn + c * plane; }
                else { p.in[c] = W + c * plane; p.out[c] = W + c * plane; }
            }
            p.ulo0 = I.abox.lo[0]; p.ulo1 = I.abox.lo[1]; p.unx = I.unx();
            p.pbytes = (unsigned)(plane * sizeof(T));
            if (onf) {
                p.bin = I.b[cur];
                p.bout = I.b[1 - cur];
                p.blo0 = I.blo0;
                p.bnx = I.bnx;
                p.bnb = L->ctx->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
            }
            p.vlo0 = v.lo[0]; p.vlo1 = v.lo[1]; p.vhi0 = v.hi[0]; p.vhi1 = v.hi[1];
            p.scen = 0;
        }
    }
}

// the islands of a box list (see Island)
static void build_islands(const std::vector<tmk_box>& boxes, const tmk_box& dom, std::vector<Island>& isl,
                          std::vector<int>& box_isl)
{
    const int n = (int)boxes.size();
    std::vector<int> par(n);
    for (int k = 0; k < n; ++k) par[k] = k;
    std::function<int(int)> root = [&](int k) { return par[k] == k ? k : (par[k] = root(par[k])); };
    // boxes whose ghost regions reach each other: sort by x and sweep
    std::vector<int> ord(n);
    for (int k = 0; k < n; ++k) ord[k] = k;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return boxes[a].lo[0] < boxes[b].lo[0]; });
    for (int p = 0; p < n; ++p) {
        const tmk_box g = box_grow(boxes[ord[p]], TMK_GHOST);
        for (int q = p + 1; q < n && boxes[ord[q]].lo[0] <= g.hi[0]; ++q)
            if (!box_empty(box_and(g, boxes[ord[q]]))) par[root(ord[q])] = root(ord[p]);
    }
    std::vector<int> comp_of(n, -1);
    std::vector<std::vector<int>> comps;
    for (int k = 0; k < n; ++k) {
        const int r = root(k);
        if (comp_of[r] < 0) {
            comp_of[r] = (int)comps.size();
            comps.emplace_back();
        }
        comps[comp_of[r]].push_back(k);
    }
    isl.clear();
    box_isl.assign(n, -1);
    auto area = [](const tmk_box& b) { return (double)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1); };
    auto add = [&](const std::vector<int>& members) {
        Island I;
        I.boxes = members;
        I.bbox = boxes[members[0]];
        for (int k : members) {
            const tmk_box& b = boxes[k];
            I.bbox.lo[0] = std::min(I.bbox.lo[0], b.lo[0]); I.bbox.lo[1] = std::min(I.bbox.lo[1], b.lo[1]);
            I.bbox.hi[0] = std::max(I.bbox.hi[0], b.hi[0]); I.bbox.hi[1] = std::max(I.bbox.hi[1], b.hi[1]);
            if (b.lo[1] == dom.lo[1]) I.has_b = 1;
            box_isl[k] = (int)isl.size();
        }
        I.abox = box_grow(I.bbox, TMK_GHOST);
        I.blo0 = I.bbox.lo[0] - TMK_GHOST;
        I.bnx = I.bbox.hi[0] - I.bbox.lo[0] + 1 + 2 * TMK_GHOST;
        isl.push_back(I);
    };
    for (const auto& cm : comps) {
        tmk_box bb = boxes[cm[0]];
        double own = 0.0;
        for (int k : cm) {
            bb.lo[0] = std::min(bb.lo[0], boxes[k].lo[0]); bb.lo[1] = std::min(bb.lo[1], boxes[k].lo[1]);
            bb.hi[0] = std::max(bb.hi[0], boxes[k].hi[0]); bb.hi[1] = std::max(bb.hi[1], boxes[k].hi[1]);
            own += area(box_grow(boxes[k], TMK_GHOST));
        }
        if (cm.size() > 1 && area(box_grow(bb, TMK_GHOST)) > 2.0 * own) {
            for (int k : cm) add(std::vector<int>{k});
        } else {
            add(cm);
        }
    }
}

// the island's boxes merged into larger rectangles (rows of boxes with equal y extent, then
// columns of those with equal x extent), tiled as the update's work units
static void island_rects(const std::vector<tmk_box>& boxes, const Island& I, std::vector<tmk_box>& out)
{
    std::vector<tmk_box> r;
    for (int k : I.boxes) r.push_back(boxes[k]);
    auto merge = [&](int d) {
        const int o = 1 - d;  // merge along d: equal extent in o, adjacent in d
        std::sort(r.begin(), r.end(), [&](const tmk_box& a, const tmk_box& b) {
            if (a.lo[o] != b.lo[o]) return a.lo[o] < b.lo[o];
            if (a.hi[o] != b.hi[o]) return a.hi[o] < b.hi[o];
            return a.lo[d] < b.lo[d];
        });
        std::vector<tmk_box> m;
        for (const tmk_box& b : r) {
            if (!m.empty() && m.back().lo[o] == b.lo[o] && m.back().hi[o] == b.hi[o] && m.back().hi[d] + 1 == b.lo[d])
                m.back().hi[d] = b.hi[d];
            else
                m.push_back(b);
        }
        r.swap(m);
    };
    merge(0);
    merge(1);
    out = r;
}

static size_t level_bytes(int precision, int fault_kind, tmk_box domain, int nboxes, const tmk_box* boxes)
{
    const size_t ts = precision == TMK_PREC_FP64 ? sizeof(double) : sizeof(float);
    const int nb = fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
    std::vector<Island> isl;
    std::vector<int> bi;
    build_islands(std::vector<tmk_box>(boxes, boxes + nboxes), domain, isl, bi);
    size_t tot = 0;
    for (const Island& I : isl) {
        tot += 3 * size_class(I.plane() * TMK_NCOMP * ts) + size_class(I.plane() * sizeof(int));
        if (I.has_b) tot += 2 * size_class((size_t)I.bnx * nb * sizeof(double)) + size_class(I.bnx * sizeof(int));
    }
    return tot;
}

extern "C" int tmk_level_footprint(const tmk_ctx* c, int nboxes, const tmk_box* boxes, size_t* bytes)
{
    if (!c || nboxes <= 0 || !boxes || !bytes) return TMK_E_ARG;
    *bytes = level_bytes(c->precision, c->hp.fault_kind, c->hp.domain, nboxes, boxes);
    return TMK_OK;
}

extern "C" int tmk_level_footprint_params(const tmk_params* p, int precision, int nboxes, const tmk_box* boxes,
                                          size_t* bytes)
{
    if (!p || nboxes <= 0 || !boxes || !bytes) return TMK_E_ARG;
    if (precision != TMK_PREC_FP64 && precision != TMK_PREC_FP32_BULK) return TMK_E_ARG;
    *bytes = level_bytes(precision, p->fault_kind, p->domain, nboxes, boxes);
    return TMK_OK;
}

static void level_free_arrays(tmk_level* L)
{
    for (Island& I : L->isl) {
        for (int k = 0; k < 2; ++k) {
            dev_free(I.u[k]);
            dev_free(I.b[k]);
            I.u[k] = nullptr;
            I.b[k] = nullptr;
        }
        dev_free(I.w);
        dev_free(I.owner);
        dev_free(I.bowner);
        I.w = nullptr;
        I.owner = I.bowner = nullptr;
    }
}

// tiles of the level's tile shape over each island's merged rectangles, and the rectangles
// themselves (the row-marching partition's input); both use the island's descriptor
static void level_build_tiles(tmk_level* L)
{
    const int nboxes = (int)L->boxes.size();
    L->tiles.clear();
    L->urects.clear();
    for (size_t k = 0; k < L->isl.size(); ++k) {
        std::vector<tmk_box> rr;
        island_rects(L->boxes, L->isl[k], rr);
        for (const tmk_box& r : rr) {
            make_tiles(r, nboxes + (int)k, L->bx - 2, L->by - 2, L->tiles);
            Tile u;
            u.patch = nboxes + (int)k;
            u.x0 = r.lo[0];
            u.y0 = r.lo[1];
            u.x1 = r.hi[0];
            u.y1 = r.hi[1];
            L->urects.push_back(u);
        }
    }
}

// Owner maps built on the device: after a fill with -1 (all bytes 0xff), workgroup row y of a
// launch writes box id[y] over its rectangle, cell (i, j) at (i - ax) + (j - ay) * unx.  The
// boxes travel in the kernel arguments, so creating a level copies nothing from the host.
constexpr int kOwnerBoxes = 32;
struct OwnerBoxes {
    int lo0[kOwnerBoxes], lo1[kOwnerBoxes], w[kOwnerBoxes], h[kOwnerBoxes], id[kOwnerBoxes];
};

static __global__ void k_owner_boxes(int* owner, int unx, int ax, int ay, OwnerBoxes ob)
{
    const int y = blockIdx.y;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int w = ob.w[y];
    if (t >= (long long)w * ob.h[y]) return;
    const int i = ob.lo0[y] + (int)(t % w), j = ob.lo1[y] + (int)(t / w);
    owner[(size_t)(i - ax) + (size_t)(j - ay) * unx] = ob.id[y];
}

// the rectangles (lo0, lo1, w, h) of boxes id into owner, kOwnerBoxes per launch
static int owner_launch(hipStream_t st, int* owner, int unx, int ax, int ay,
                        const std::vector<std::array<int, 5>>& rects)
{
    for (size_t k0 = 0; k0 < rects.size(); k0 += kOwnerBoxes) {
        OwnerBoxes ob;
        memset(&ob, 0, sizeof ob);
        const int n = (int)std::min<size_t>(kOwnerBoxes, rects.size() - k0);
        long long most = 1;
        for (int y = 0; y < n; ++y) {
            const std::array<int, 5>& r = rects[k0 + y];
            ob.lo0[y] = r[0];
            ob.lo1[y] = r[1];
            ob.w[y] = r[2];
            ob.h[y] = r[3];
            ob.id[y] = r[4];
            most = std::max(most, (long long)r[2] * r[3]);
        }
        hipLaunchKernelGGL(k_owner_boxes, dim3((unsigned)((most + 255) / 256), (unsigned)n), dim3(256), 0, st, owner,
                           unx, ax, ay, ob);
        HIPCHK(hipGetLastError());
    }
    return TMK_OK;
}

extern "C" int tmk_level_create(tmk_ctx* c, int nboxes, const tmk_box* boxes, tmk_level** out)
{
    if (!c || !out || nboxes <= 0 || !boxes) return TMK_E_ARG;
    *out = nullptr;
    const tmk_box dom = c->hp.domain;
    for (int b = 0; b < nboxes; ++b)
        if (box_empty(boxes[b]) || !box_contains(dom, boxes[b])) return TMK_E_BOX;
    tmk_level* L = new (std::nothrow) tmk_level;
    if (!L) return TMK_E_NOMEM;
    L->ctx = c;
    L->boxes.assign(boxes, boxes + nboxes);
    L->bx = c->bx;
    L->by = c->by;
    L->march = c->march;
    L->ns = c->antiplane ? 3 : 2;
    L->tsize = c->precision == TMK_PREC_FP64 ? sizeof(double) : sizeof(float);
    L->bytes = level_bytes(c->precision, c->hp.fault_kind, c->hp.domain, nboxes, boxes);
    const int nb = c->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
    for (int b = 0; b < nboxes; ++b) L->has_b.push_back(boxes[b].lo[1] == dom.lo[1]);
    build_islands(L->boxes, dom, L->isl, L->box_isl);
    bool ok = true;
    for (Island& I : L->isl) {
        const size_t plane = I.plane();
        const size_t bytes = plane * TMK_NCOMP * L->tsize;
        // the kernels address a plane with 32-bit byte offsets (gld/gst)
        ok = ok && plane * L->tsize <= kPlaneMaxBytes;
        ok = ok && dev_alloc(&I.u[0], bytes) == hipSuccess && dev_alloc(&I.u[1], bytes) == hipSuccess &&
             dev_alloc(&I.w, bytes) == hipSuccess && dev_alloc(&I.owner, plane * sizeof(int)) == hipSuccess;
        // owner maps, enqueued on the context's stream: every use of the level is on that stream
        // or synchronises it first (batch and hierarchy membership, regrid, destroy)
        if (ok) {
            std::vector<std::array<int, 5>> rects;
            for (int k : I.boxes) {
                const tmk_box v = L->boxes[k];
                rects.push_back({v.lo[0], v.lo[1], v.hi[0] - v.lo[0] + 1, v.hi[1] - v.lo[1] + 1, k});
            }
            ok = hipMemsetAsync(I.owner, 0xff, plane * sizeof(int), c->stream) == hipSuccess &&
                 owner_launch(c->stream, I.owner, I.unx(), I.abox.lo[0], I.abox.lo[1], rects) == TMK_OK;
        }
        if (ok && I.has_b) {
            const size_t nB = (size_t)I.bnx * nb;
            ok = dev_alloc(&I.b[0], nB * sizeof(double)) == hipSuccess &&
                 dev_alloc(&I.b[1], nB * sizeof(double)) == hipSuccess &&
                 dev_alloc(&I.bowner, I.bnx * sizeof(int)) == hipSuccess;
            if (ok) {
                std::vector<std::array<int, 5>> rects;
                for (int k : I.boxes)
                    if (L->has_b[k]) rects.push_back({L->boxes[k].lo[0], 0, L->boxes[k].hi[0] - L->boxes[k].lo[0] + 1, 1, k});
                ok = hipMemsetAsync(I.bowner, 0xff, I.bnx * sizeof(int), c->stream) == hipSuccess &&
                     owner_launch(c->stream, I.bowner, I.bnx, I.blo0, 0, rects) == TMK_OK;
            }
        }
        if (!ok) break;
    }
    if (!ok) {
        level_free_arrays(L);
        delete L;
        return TMK_E_NOMEM;
    }
    if (L->tsize == sizeof(double)) level_fill_patches<double>(L);
    else level_fill_patches<float>(L);
    level_build_tiles(L);
    // exchange plan: ghost cells (inside the domain) of each box covered by valid cells of a
    // box in another island
    for (int d = 0; d < nboxes; ++d) {
        const tmk_box g = box_and(box_grow(boxes[d], TMK_GHOST), dom);
        for (int s = 0; s < nboxes; ++s) {
            if (s == d || L->box_isl[s] == L->box_isl[d]) continue;
            const tmk_box r = box_and(g, boxes[s]);
            if (box_empty(r)) continue;
            L->rects.push_back(CopyRect{s, d, r.lo[0], r.lo[1], r.hi[0], r.hi[1]});
            if (L->has_b[d] && L->has_b[s]) {
                const int lo = std::max(boxes[d].lo[0] - TMK_GHOST, boxes[s].lo[0]);
                const int hi = std::min(boxes[d].hi[0] + TMK_GHOST, boxes[s].hi[0]);
                if (lo <= hi) L->brects.push_back(CopyRect{s, d, lo, 0, hi, 0});
            }
        }
    }
    *out = L;
    return TMK_OK;
}

extern "C" int tmk_level_destroy(tmk_level* L)
{
    if (!L) return TMK_OK;
    if (L->batch || L->amr_stream) return TMK_E_ARG;
    if (L->ctx && L->ctx->stream) (void)hipStreamSynchronize(L->ctx->stream);
    if (L->stations) L->stations->L = nullptr;
    L->own.release();
    level_free_arrays(L);
    delete L;
    return TMK_OK;
}

// a box's ghosted region <-> staging array (22 planes of the ghosted box, x fastest); put
// writes the cells the box owns and the island's ghost cells, never another box's cells
template <typename T>
static __global__ void k_box_put(T* u0, T* u1, T* w, const int* owner, int aunx, size_t aplane, int ox, int oy,
                                 const T* src, int gnx, int gny, int box)
{
    const size_t n = (size_t)gnx * gny;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int x = (int)(t % gnx), y = (int)(t / gnx);
    const size_t a = (size_t)(x + ox) + (size_t)(y + oy) * aunx;
    const int o = owner[a];
    if (o != box && o != -1) return;
    for (int c = 0; c < TMK_NCOMP; ++c) {
        const T v = src[t + c * n];
        u0[a + c * aplane] = v;
        u1[a + c * aplane] = v;
        w[a + c * aplane] = v;
    }
}

template <typename T>
static __global__ void k_box_const(T* u0, T* u1, T* w, const int* owner, int aunx, size_t aplane, int ox, int oy,
                                   StateVec st, int gnx, int gny, int box)
{
    const size_t n = (size_t)gnx * gny;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int x = (int)(t % gnx), y = (int)(t / gnx);
    const size_t a = (size_t)(x + ox) + (size_t)(y + oy) * aunx;
    const int o = owner[a];
    if (o != box && o != -1) return;
    for (int c = 0; c < TMK_NCOMP; ++c) {
        const T v = (T)st.c[c];
        u0[a + c * aplane] = v;
        u1[a + c * aplane] = v;
        w[a + c * aplane] = v;
    }
}

template <typename T>
static __global__ void k_box_get(const T* u, const T* w, int aunx, size_t aplane, int ox, int oy, T* dst, int gnx,
                                 int gny, int smask)
{
    const size_t n = (size_t)gnx * gny;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int x = (int)(t % gnx), y = (int)(t / gnx);
    const size_t a = (size_t)(x + ox) + (size_t)(y + oy) * aunx;
    for (int c = 0; c < TMK_NCOMP; ++c) dst[t + c * n] = ((smask >> c) & 1) ? u[a + c * aplane] : w[a + c * aplane];
}

// fault-data row of a box (with its ghosts) <-> the island's row
static __global__ void k_row_put(double* b0, double* b1, const int* bowner, int off, int bnx, const double* src, int n,
                                 int nb, int box)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    const int a = off + f;
    const int o = bowner[a];
    if (o != box && o != -1) return;
    for (int c = 0; c < nb; ++c) {
        b0[a + (size_t)c * bnx] = src[f + (size_t)c * n];
        b1[a + (size_t)c * bnx] = src[f + (size_t)c * n];
    }
}

static __global__ void k_row_get(const double* b, int off, int bnx, double* dst, int n, int nb)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= n) return;
    for (int c = 0; c < nb; ++c) dst[f + (size_t)c * n] = b[off + f + (size_t)c * bnx];
}

static int level_put_fault(tmk_level* L, int b, const double* B)
{
    tmk_ctx* c = L->ctx;
    const Island& I = L->isl[L->box_isl[b]];
    const int nb = c->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
    const int n = L->boxes[b].hi[0] - L->boxes[b].lo[0] + 1 + 2 * TMK_GHOST;
    int e;
    if ((e = c->bufs[0].ensure((size_t)n * nb * sizeof(double)))) return e;
    HIPCHK(copy_on(c->stream, c->bufs[0].p, B, (size_t)n * nb * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_row_put, dim3((n + 255) / 256), dim3(256), 0, c->stream, I.b[0], I.b[1], I.bowner,
                       L->boxes[b].lo[0] - TMK_GHOST - I.blo0, I.bnx, (const double*)c->bufs[0].p, n, nb, b);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return TMK_OK;
}

static int level_get_fault(tmk_level* L, int b, double* B, hipStream_t st)
{
    tmk_ctx* c = L->ctx;
    const Island& I = L->isl[L->box_isl[b]];
    const int nb = c->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
    const int n = L->boxes[b].hi[0] - L->boxes[b].lo[0] + 1 + 2 * TMK_GHOST;
    int e;
    if ((e = c->bufs[0].ensure((size_t)n * nb * sizeof(double)))) return e;
    hipLaunchKernelGGL(k_row_get, dim3((n + 255) / 256), dim3(256), 0, st, I.b[L->cur],
                       L->boxes[b].lo[0] - TMK_GHOST - I.blo0, I.bnx, (double*)c->bufs[0].p, n, nb);
    HIPCHK(hipGetLastError());
    HIPCHK(copy_on(st, B, c->bufs[0].p, (size_t)n * nb * sizeof(double), hipMemcpyDeviceToHost));
    return TMK_OK;
}

extern "C" int tmk_level_upload(tmk_level* L, int b, const double* U, const double* B)
{
    if (!L || b < 0 || b >= (int)L->boxes.size() || !U || L->batch || L->amr_stream) return TMK_E_ARG;
    tmk_ctx* c = L->ctx;
    const tmk_box ub = box_grow(L->boxes[b], TMK_GHOST);
    const int gnx = ub.hi[0] - ub.lo[0] + 1, gny = ub.hi[1] - ub.lo[1] + 1;
    const size_t plane = (size_t)gnx * gny;
    const size_t nU = plane * TMK_NCOMP;
    if (L->ns == 2) {  // mode II level: the anti-plane planes must be zero
        for (int s = 0; s < 2; ++s)
            for (int cc : {2, 7, 8})
                for (size_t k = 0; k < plane; ++k)
                    if (U[k + (size_t)(11 * s + cc) * plane] != 0.0) return TMK_E_ARG;
    }
    if (L->has_b[b] && !B) return TMK_E_NO_BDRY;
    std::vector<float> tmp;
    const void* src = U;
    if (L->tsize == sizeof(float)) {
        tmp.resize(nU);
        to_stored_f32(U, tmp.data(), plane, c->P);
        src = tmp.data();
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    int e;
    if ((e = c->bufs[7].ensure(nU * L->tsize))) return e;
    HIPCHK(copy_on(c->stream, c->bufs[7].p, src, nU * L->tsize, hipMemcpyHostToDevice));
    const Island& I = L->isl[L->box_isl[b]];
    const int ox = ub.lo[0] - I.abox.lo[0], oy = ub.lo[1] - I.abox.lo[1];
    const dim3 grid((unsigned)((plane + 255) / 256));
    if (L->tsize == sizeof(double))
        hipLaunchKernelGGL(k_box_put<double>, grid, dim3(256), 0, c->stream, (double*)I.u[0], (double*)I.u[1],
                           (double*)I.w, I.owner, I.unx(), I.plane(), ox, oy, (const double*)c->bufs[7].p, gnx, gny, b);
    else
        hipLaunchKernelGGL(k_box_put<float>, grid, dim3(256), 0, c->stream, (float*)I.u[0], (float*)I.u[1],
                           (float*)I.w, I.owner, I.unx(), I.plane(), ox, oy, (const float*)c->bufs[7].p, gnx, gny, b);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    if (L->has_b[b] && (e = level_put_fault(L, b, B))) return e;
    L->dirty = true;
    return TMK_OK;
}

extern "C" int tmk_level_fill(tmk_level* L, int b, const double* state, const double* B)
{
    if (!L || b < 0 || b >= (int)L->boxes.size() || !state || L->batch || L->amr_stream) return TMK_E_ARG;
    if (L->ns == 2 && (state[2] != 0.0 || state[7] != 0.0 || state[8] != 0.0 || state[13] != 0.0 ||
                       state[18] != 0.0 || state[19] != 0.0))
        return TMK_E_ARG;
    if (L->has_b[b] && !B) return TMK_E_NO_BDRY;
    tmk_ctx* c = L->ctx;
    const tmk_box ub = box_grow(L->boxes[b], TMK_GHOST);
    const int gnx = ub.hi[0] - ub.lo[0] + 1, gny = ub.hi[1] - ub.lo[1] + 1;
    const size_t plane = (size_t)gnx * gny;
    StateVec v;
    for (int k = 0; k < TMK_NCOMP; ++k) v.c[k] = L->tsize == sizeof(float) ? state[k] - c->P.bg[k] : state[k];
    const Island& I = L->isl[L->box_isl[b]];
    const int ox = ub.lo[0] - I.abox.lo[0], oy = ub.lo[1] - I.abox.lo[1];
    const dim3 grid((unsigned)((plane + 255) / 256));
    if (L->tsize == sizeof(double))
        hipLaunchKernelGGL(k_box_const<double>, grid, dim3(256), 0, c->stream, (double*)I.u[0], (double*)I.u[1],
                           (double*)I.w, I.owner, I.unx(), I.plane(), ox, oy, v, gnx, gny, b);
    else
        hipLaunchKernelGGL(k_box_const<float>, grid, dim3(256), 0, c->stream, (float*)I.u[0], (float*)I.u[1],
                           (float*)I.w, I.owner, I.unx(), I.plane(), ox, oy, v, gnx, gny, b);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    int e;
    if (L->has_b[b] && (e = level_put_fault(L, b, B))) return e;
    L->dirty = true;
    return TMK_OK;
}

/* fault data only (no state) of one box, e.g. for result records of a running scenario */
extern "C" int tmk_level_download_fault(tmk_level* L, int b, double* B)
{
    if (!L || b < 0 || b >= (int)L->boxes.size() || !B || L->batch) return TMK_E_ARG;
    if (!L->has_b[b]) return TMK_E_NO_BDRY;
    hipStream_t st = L->amr_stream ? L->amr_stream : L->ctx->stream;
    HIPCHK(hipStreamSynchronize(st));
    return level_get_fault(L, b, B, st);
}

extern "C" int tmk_level_download(tmk_level* L, int b, double* U, double* B)
{
    if (!L || b < 0 || b >= (int)L->boxes.size() || !U || L->batch) return TMK_E_ARG;
    tmk_ctx* c = L->ctx;
    hipStream_t st = L->amr_stream ? L->amr_stream : c->stream;
    HIPCHK(hipStreamSynchronize(st));
    const tmk_box ub = box_grow(L->boxes[b], TMK_GHOST);
    const int gnx = ub.hi[0] - ub.lo[0] + 1, gny = ub.hi[1] - ub.lo[1] + 1;
    const size_t plane = (size_t)gnx * gny;
    const size_t nU = plane * TMK_NCOMP;
    int e;
    if ((e = c->bufs[7].ensure(nU * L->tsize))) return e;
    const Island& I = L->isl[L->box_isl[b]];
    const int ox = ub.lo[0] - I.abox.lo[0], oy = ub.lo[1] - I.abox.lo[1];
    const dim3 grid((unsigned)((plane + 255) / 256));
    if (L->tsize == sizeof(double))
        hipLaunchKernelGGL(k_box_get<double>, grid, dim3(256), 0, st, (const double*)I.u[L->cur], (const double*)I.w,
                           I.unx(), I.plane(), ox, oy, (double*)c->bufs[7].p, gnx, gny, stencil_mask(L->ns));
    else
        hipLaunchKernelGGL(k_box_get<float>, grid, dim3(256), 0, st, (const float*)I.u[L->cur], (const float*)I.w,
                           I.unx(), I.plane(), ox, oy, (float*)c->bufs[7].p, gnx, gny, stencil_mask(L->ns));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    if (L->tsize == sizeof(double)) {
        HIPCHK(copy_on(st, U, c->bufs[7].p, nU * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> a(nU);
        HIPCHK(copy_on(st, a.data(), c->bufs[7].p, nU * sizeof(float), hipMemcpyDeviceToHost));
        from_stored_f32(a.data(), U, plane, c->P);
    }
    if (L->has_b[b] && B) return level_get_fault(L, b, B, st);
    return TMK_OK;
}

// make the level's own one-member group current (built lazily, rebuilt when the
// level left a batch at another parity)
static int level_own_group(tmk_level* L)
{
    if (!L->own_built) {
        std::vector<tmk_level*> m{L};
        std::vector<int> ph{L->cur};
        // the rebuild rewrites the group's device tables in place
        HIPCHK(hipStreamSynchronize(L->ctx->stream));
        int e = group_build(L->own, m, ph, 1);
        if (e) return e;
        L->own.stations = L->stations ? &L->stations->dev : nullptr;
        L->own_built = true;
    }
    return TMK_OK;
}

static int level_prepare(tmk_level* L, double time, double dt)
{
    int e = level_own_group(L);
    if (e) return e;
    hipStream_t st = L->ctx->stream;
    // group parity tracks the level parity (phase 0 at build time == cur at build)
    if (L->dirty) {
        L->dirty = false;
        e = L->tsize == sizeof(double) ? group_exchange_now_t<double>(L->own, st)
                                       : group_exchange_now_t<float>(L->own, st);
        if (e) return e;
    }
    // the clock lives in device memory, so captured graphs stay valid across calls
    L->time = time;
    L->dt = dt;
    return group_set_clock(L->own, st, 0, time, dt, L->ctx->P);
}

extern "C" int tmk_level_step(tmk_level* L, double time, double dt, double* dt_new)
{
    if (!L || L->batch || L->amr_stream) return TMK_E_ARG;
    int e = level_prepare(L, time, dt);
    if (e) return e;
    if ((e = group_enqueue_step(L->own, L->ctx->stream))) return e;
    L->cur ^= 1;
    L->time = time + dt;
    if (dt_new) *dt_new = L->ctx->hp.dx / std::max(L->ctx->hp.cs, L->ctx->hp.cp);
    return TMK_OK;
}

extern "C" int tmk_level_run(tmk_level* L, int nsteps, double time, double dt)
{
    if (!L || nsteps < 0 || L->batch || L->amr_stream) return TMK_E_ARG;
    if (nsteps == 0) return TMK_OK;
    int e = level_prepare(L, time, dt);
    if (e) return e;
    if ((e = group_run(L->own, L->ctx->stream, nsteps))) return e;
    L->cur ^= (nsteps & 1);
    double t = time;
    for (int k = 0; k < nsteps; ++k) t = t + dt;
    L->time = t;
    return TMK_OK;
}

extern "C" int tmk_level_run_timed(tmk_level* L, int nsteps, double time, double dt, double* ms)
{
    if (!L || !ms) return TMK_E_ARG;
    hipStream_t st = L->ctx->stream;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    int e = level_prepare(L, time, dt);
    if (!e && hipEventRecord(a, st) != hipSuccess) e = TMK_E_HIP;
    if (!e) e = tmk_level_run(L, nsteps, time, dt);
    if (!e && hipEventRecord(b, st) != hipSuccess) e = TMK_E_HIP;
    if (!e && hipEventSynchronize(b) != hipSuccess) e = TMK_E_HIP;
    float f = 0.f;
    if (!e && hipEventElapsedTime(&f, a, b) != hipSuccess) e = TMK_E_HIP;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!e) *ms = f;
    return e;
}

extern "C" int tmk_level_profile(tmk_level* L, int nsteps, double time, double dt, double* ms)
{
    if (!L || !ms || nsteps <= 0 || L->batch || L->amr_stream) return TMK_E_ARG;
    int e = level_prepare(L, time, dt);
    if (e) return e;
    hipStream_t st = L->ctx->stream;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    if (hipEventRecord(a, st) != hipSuccess) e = TMK_E_HIP;
    for (int k = 0; k < nsteps && !e; ++k) {
        e = group_enqueue_step(L->own, st);
        L->cur ^= 1;
        L->time = L->time + dt;
    }
    if (!e && hipEventRecord(b, st) != hipSuccess) e = TMK_E_HIP;
    if (!e && hipEventSynchronize(b) != hipSuccess) e = TMK_E_HIP;
    float f = 0.f;
    if (!e && hipEventElapsedTime(&f, a, b) != hipSuccess) e = TMK_E_HIP;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!e) *ms = (double)f / nsteps;
    return e;
}

// nsteps steps of a level with timing events around each step and each update launch; ev holds
// 6 events per step.  Returns the error of the enqueue.
static int profile_enqueue(tmk_level* L, hipStream_t st, int nsteps, std::vector<hipEvent_t>& ev)
{
    for (int k = 0; k < nsteps; ++k) {
        hipEvent_t* x = &ev[(size_t)6 * k];
        StepTimers tm;
        tm.ia = x[2]; tm.ib = x[3]; tm.ba = x[4]; tm.bb = x[5];
        if (hipEventRecord(x[0], st) != hipSuccess) return TMK_E_HIP;
        int e = group_enqueue_step(L->own, st, &tm);
        if (e) return e;
        if (hipEventRecord(x[1], st) != hipSuccess) return TMK_E_HIP;
    }
    return TMK_OK;
}

// the same steps with wall-clock stamps instead of events (6 per step: the four launch marks,
// then the step's start and end on the level's stream)
static int profile_enqueue_stamps(tmk_level* L, hipStream_t st, int nsteps, long long* stamps)
{
    for (int k = 0; k < nsteps; ++k) {
        StepTimers tm;
        tm.stamps = stamps + (size_t)6 * k;
        hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, st, tm.stamps + 4);
        if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
        int e = group_enqueue_step(L->own, st, &tm);
        if (e) return e;
        hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, st, tm.stamps + 5);
        if (hipGetLastError() != hipSuccess) return TMK_E_HIP;
    }
    return TMK_OK;
}

extern "C" int tmk_level_profile_kernels(tmk_level* L, int nsteps, double time, double dt, tmk_profile* out)
{
    if (!L || !out || nsteps <= 0 || L->batch || L->amr_stream) return TMK_E_ARG;
    int e = level_prepare(L, time, dt);
    if (e) return e;
    hipStream_t st = L->ctx->stream;
    Group& G = L->own;
    const bool has_bnd = G.ntiles > G.nint, has_int = G.nint > 0;
    double sum[3] = {0, 0, 0};
    // The steps as one captured graph, replayed once, so that the launches run as in
    // tmk_level_run; each mark is a one-lane kernel reading the constant-rate wall clock.
    // Where the clock rate is unknown or capture fails, the same steps run eagerly with events.
    int in_graph = 0;
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess &&
        khz > 0) {
        DevBuf sb;
        const size_t ns = (size_t)6 * nsteps;
        if ((e = sb.ensure(ns * sizeof(long long)))) return e;
        const int g0 = G.g;
        hipGraph_t gr = nullptr;
        hipGraphExec_t gx = nullptr;
        bool ran = false;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int ce = profile_enqueue_stamps(L, st, nsteps, (long long*)sb.p);
        const hipError_t ee = hipStreamEndCapture(st, &gr);
        G.g = g0;  // nothing has run yet
        if (!ce && ee == hipSuccess && hipGraphInstantiate(&gx, gr, nullptr, nullptr, 0) == hipSuccess &&
            hipGraphLaunch(gx, st) == hipSuccess)
            ran = true;
        if (gx) (void)hipGraphExecDestroy(gx);
        if (gr) (void)hipGraphDestroy(gr);
        if (ce) return ce;
        if (ran) {
            std::vector<long long> h(ns);
            HIPCHK(hipMemcpyAsync(h.data(), sb.p, ns * sizeof(long long), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            for (int k = 0; k < nsteps; ++k) {
                G.g ^= 1;
                L->cur ^= 1;
                L->time = L->time + dt;
                const long long* x = &h[(size_t)6 * k];
                // the wall clock counts up and wraps as an unsigned 64-bit counter
                auto ticks = [](long long a, long long b) {
                    return (double)(long long)((unsigned long long)b - (unsigned long long)a);
                };
                sum[0] += ticks(x[4], x[5]) / khz;  // kHz: ticks per ms
                if (has_int) sum[1] += ticks(x[0], x[1]) / khz;
                if (has_bnd) sum[2] += ticks(x[2], x[3]) / khz;
            }
            in_graph = 1;
        }
        sb.release();
    }
    if (!in_graph) {
        std::vector<hipEvent_t> ev((size_t)6 * nsteps, nullptr);
        for (auto& x : ev)
            if (hipEventCreate(&x) != hipSuccess) e = TMK_E_HIP;
        if (!e) e = profile_enqueue(L, st, nsteps, ev);
        for (int k = 0; k < nsteps; ++k) {
            L->cur ^= 1;
            L->time = L->time + dt;
        }
        if (!e && hipStreamSynchronize(st) != hipSuccess) e = TMK_E_HIP;
        for (int k = 0; k < nsteps && !e; ++k) {
            const hipEvent_t* x = &ev[(size_t)6 * k];
            float f = 0.f;
            if (hipEventElapsedTime(&f, x[0], x[1]) != hipSuccess) { e = TMK_E_HIP; break; }
            sum[0] += f;
            if (has_int) {
                if (hipEventElapsedTime(&f, x[2], x[3]) != hipSuccess) { e = TMK_E_HIP; break; }
                sum[1] += f;
            }
            if (has_bnd) {
                if (hipEventElapsedTime(&f, x[4], x[5]) != hipSuccess) { e = TMK_E_HIP; break; }
                sum[2] += f;
            }
        }
        for (auto& x : ev)
            if (x) (void)hipEventDestroy(x);
        if (e) return e;
    }
    out->ms_step = sum[0] / nsteps;
    out->ms_interior = sum[1] / nsteps;
    out->ms_boundary = sum[2] / nsteps;
    out->cells_interior = G.cells_int;
    out->cells_boundary = G.cells_bnd;
    out->tiles_interior = G.nint;
    out->tiles_boundary = G.ntiles - G.nint;
    out->in_graph = in_graph;
    return TMK_OK;
}

extern "C" int tmk_level_sync(tmk_level* L)
{
    if (!L || L->batch || L->amr_stream) return TMK_E_ARG;
    HIPCHK(hipStreamSynchronize(L->ctx->stream));
    if (!L->own_built) return TMK_OK;
    int code = 0;
    int e = group_error(L->own, L->ctx->stream, 0, &code);
    if (e) return e;
    if (code) {
        const int z = 0;
        HIPCHK(copy_on(L->ctx->stream, (char*)(L->own.scen_dev) + offsetof(Scen, err), &z, sizeof z,
                       hipMemcpyHostToDevice));
    }
    return code;
}

// ---------------------------------------------------------------------------
// update shape of a level (row-marching strips or tiles) and its tuning
// ---------------------------------------------------------------------------
extern "C" int tmk_level_set_update_shape(tmk_level* L, int march_rows, int bx, int by)
{
    if (!L || L->batch || L->amr_stream || march_rows < 0 || march_rows > 65536) return TMK_E_ARG;
    if (bx == 0 && by == 0) default_tile(L->ctx->precision, bx, by);
    if (!tile_ok(bx, by)) return TMK_E_ARG;
    HIPCHK(hipStreamSynchronize(L->ctx->stream));
    L->march = march_rows;
    if (bx != L->bx || by != L->by) {
        L->bx = bx;
        L->by = by;
        level_build_tiles(L);
    }
    L->own_built = false;  // tables and graph are rebuilt at the next step
    return TMK_OK;
}

extern "C" int tmk_level_set_tile_order(tmk_level* L, int xcd)
{
    if (!L || L->batch || L->amr_stream || (xcd != 0 && xcd != 1)) return TMK_E_ARG;
    HIPCHK(hipStreamSynchronize(L->ctx->stream));
    L->xcd = xcd;
    L->own_built = false;
    return TMK_OK;
}

extern "C" int tmk_level_get_update_shape(const tmk_level* L, int* march_rows, int* bx, int* by)
{
    if (!L || !march_rows || !bx || !by) return TMK_E_ARG;
    const bool m = L->march > 0 && L->ns == 2 && pred_kind(L->ctx->P) == PK_PPM_PRIM;
    *march_rows = m ? L->march : 0;
    *bx = L->bx;
    *by = L->by;
    return TMK_OK;
}

// the strip heights and tile shapes the tuner times
static const int kTuneMarch[] = {4, 6, 8, 12, 16, 20, 24, 28, 32, 40, 48, 64};
static const int kTuneTile[][2] = {{32, 16}, {32, 8}, {32, 12}, {32, 32}, {16, 16}, {64, 4}, {64, 8}, {64, 16}};

// a device copy of a level's state (every island's current stencil planes, in-place planes and
// fault rows), to run tuning steps from and return to
struct LevelSnapshot {
    std::vector<void*> bufs;
    int cur = 0;
    ~LevelSnapshot()
    {
        for (void* p : bufs)
            if (p) (void)hipFree(p);
    }
    static size_t fault_bytes(const tmk_level* L, const Island& I)
    {
        const int nb = L->ctx->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
        return I.has_b ? (size_t)I.bnx * nb * sizeof(double) : 0;
    }
    int take(const tmk_level* L, hipStream_t st)
    {
        cur = L->cur;
        for (const Island& I : L->isl) {
            const size_t sb = I.plane() * TMK_NCOMP * L->tsize, fb = fault_bytes(L, I);
            void *u = nullptr, *w = nullptr, *b = nullptr;
            if (hipMalloc(&u, sb) != hipSuccess || hipMalloc(&w, sb) != hipSuccess ||
                (fb && hipMalloc(&b, fb) != hipSuccess)) {
                if (u) (void)hipFree(u);
                if (w) (void)hipFree(w);
                return TMK_E_NOMEM;
            }
            bufs.push_back(u);
            bufs.push_back(w);
            bufs.push_back(b);
            HIPCHK(hipMemcpyAsync(u, I.u[cur], sb, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(w, I.w, sb, hipMemcpyDeviceToDevice, st));
            if (fb) HIPCHK(hipMemcpyAsync(b, I.b[cur], fb, hipMemcpyDeviceToDevice, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        return TMK_OK;
    }
    int restore(tmk_level* L, hipStream_t st) const
    {
        for (size_t k = 0; k < L->isl.size(); ++k) {
            const Island& I = L->isl[k];
            const size_t sb = I.plane() * TMK_NCOMP * L->tsize, fb = fault_bytes(L, I);
            HIPCHK(hipMemcpyAsync(I.u[cur], bufs[3 * k], sb, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(I.w, bufs[3 * k + 1], sb, hipMemcpyDeviceToDevice, st));
            if (fb) HIPCHK(hipMemcpyAsync(I.b[cur], bufs[3 * k + 2], fb, hipMemcpyDeviceToDevice, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        L->cur = cur;
        L->own_built = false;  // the group's parity follows L->cur from its next build
        return TMK_OK;
    }
};

extern "C" int tmk_level_autotune(tmk_level* L, double time, double dt, int nsteps, tmk_tuning* out)
{
    if (!L || !out || nsteps < 2 || L->batch || L->amr_stream || L->stations) return TMK_E_ARG;
    tmk_ctx* c = L->ctx;
    hipStream_t st = c->stream;
    memset(out, 0, sizeof *out);
    int e = level_prepare(L, time, dt);  // exchanges uploaded data first: the snapshot is a valid state
    if (e) return e;
    HIPCHK(hipStreamSynchronize(st));
    LevelSnapshot snap;
    if ((e = snap.take(L, st))) return e;
    const int march0 = L->march, bx0 = L->bx, by0 = L->by, xcd0 = L->xcd;
    std::vector<std::array<int, 3>> cand;
    if (L->ns == 2 && pred_kind(c->P) == PK_PPM_PRIM)
        for (int rows : kTuneMarch) cand.push_back({rows, bx0, by0});
    for (const auto& t : kTuneTile) cand.push_back({0, t[0], t[1]});
    if ((int)cand.size() > TMK_TUNE_MAX) cand.resize(TMK_TUNE_MAX);
    int best = -1;
    for (size_t k = 0; k < cand.size() && !e; ++k) {
        if ((e = snap.restore(L, st))) break;
        if ((e = tmk_level_set_update_shape(L, cand[k][0], cand[k][1], cand[k][2]))) break;
        double ms = 0.0;
        // two untimed steps (group build, graph capture), then nsteps graph-replayed steps
        if ((e = tmk_level_run(L, 2, time, dt))) break;
        if ((e = tmk_level_run_timed(L, nsteps, time + dt + dt, dt, &ms))) break;
        if ((e = tmk_level_sync(L))) break;
        out->march_rows_tried[k] = cand[k][0];
        out->bx_tried[k] = cand[k][1];
        out->by_tried[k] = cand[k][2];
        out->ms_tried[k] = ms / nsteps;
        out->ntried = (int)k + 1;
        if (best < 0 || out->ms_tried[k] < out->ms_tried[best]) best = (int)k;
    }
    for (int k = 0; k < out->ntried; ++k) out->xcd_tried[k] = xcd0;
    // the order of the fastest shape's interior workgroups: the other of XCD bands and row-major
    int best_xcd = xcd0;
    if (!e && best >= 0 && out->ntried < TMK_TUNE_MAX) {
        const int k = out->ntried;
        double ms = 0.0;
        if (!(e = snap.restore(L, st)) &&
            !(e = tmk_level_set_update_shape(L, cand[best][0], cand[best][1], cand[best][2])) &&
            !(e = tmk_level_set_tile_order(L, 1 - xcd0)) && !(e = tmk_level_run(L, 2, time, dt)) &&
            !(e = tmk_level_run_timed(L, nsteps, time + dt + dt, dt, &ms)) && !(e = tmk_level_sync(L))) {
            out->march_rows_tried[k] = cand[best][0];
            out->bx_tried[k] = cand[best][1];
            out->by_tried[k] = cand[best][2];
            out->xcd_tried[k] = 1 - xcd0;
            out->ms_tried[k] = ms / nsteps;
            out->ntried = k + 1;
            if (out->ms_tried[k] < out->ms_tried[best]) {
                best = k;
                best_xcd = 1 - xcd0;
            }
        }
    }
    int e2 = snap.restore(L, st);
    L->time = time;
    if (e || best < 0) {
        (void)tmk_level_set_update_shape(L, march0, bx0, by0);
        (void)tmk_level_set_tile_order(L, xcd0);
        return e ? e : e2;
    }
    const int rows = out->march_rows_tried[best], bx = out->bx_tried[best], by = out->by_tried[best];
    if ((e = tmk_level_set_update_shape(L, rows, bx, by)) || (e = tmk_level_set_tile_order(L, best_xcd))) return e;
    out->march_rows = rows;
    out->bx = bx;
    out->by = by;
    out->xcd = best_xcd;
    out->ms_per_step = out->ms_tried[best];
    return e2;
}

// ---------------------------------------------------------------------------
// batches: boxes of many scenarios in shared launches (continuous batching)
// ---------------------------------------------------------------------------
extern "C" int tmk_batch_create(tmk_batch** out)
{
    if (!out) return TMK_E_ARG;
    tmk_batch* B = new (std::nothrow) tmk_batch;
    if (!B) return TMK_E_NOMEM;
    if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess) {
        delete B;
        return TMK_E_HIP;
    }
    *out = B;
    return TMK_OK;
}

extern "C" int tmk_batch_destroy(tmk_batch* B)
{
    if (!B) return TMK_OK;
    (void)hipStreamSynchronize(B->stream);
    for (tmk_level* L : B->members) L->batch = nullptr;
    B->grp.release();
    B->xgrp.release();
    (void)hipStreamDestroy(B->stream);
    delete B;
    return TMK_OK;
}

// pull each member's clock and parity out of the group before the layout changes
static int batch_detach_all(tmk_batch* B)
{
    HIPCHK(hipStreamSynchronize(B->stream));
    if (!B->built) return TMK_OK;
    std::vector<Scen> sc(B->members.size());
    if (!sc.empty())
        HIPCHK(copy_on(B->stream, sc.data(), B->grp.scen_dev, sizeof(Scen) * sc.size(), hipMemcpyDeviceToHost));
    for (size_t m = 0; m < B->members.size(); ++m) {
        tmk_level* L = B->members[m];
        L->cur = B->phase[m] ^ B->grp.g;
        L->time = sc[m].time;
        if (sc[m].err) L->own_built = false;
        B->phase[m] = L->cur;
    }
    B->grp.reset();
    B->built = false;
    return TMK_OK;
}

// Build the fused group for the current members.  Called lazily by the first
// run after membership changes, so adding or removing several members costs one
// rebuild.  The batch stream is idle here (batch_detach_all synchronised it).
static int batch_rebuild(tmk_batch* B)
{
    B->stale = false;
    for (size_t m = 0; m < B->members.size(); ++m) B->phase[m] = B->members[m]->cur;
    int e = group_build(B->grp, B->members, B->phase, 1);
    if (e) return e;
    B->built = !B->members.empty();
    // exchange members that were uploaded since their last step (a one-box level
    // has no same-level copies: its ghosts are filled by the boundary stencils)
    for (tmk_level* L : B->members) {
        if (!L->dirty) continue;
        L->dirty = false;
        if (L->rects.empty() && L->brects.empty()) continue;
        std::vector<tmk_level*> m{L};
        std::vector<int> ph{L->cur};
        if ((e = group_build(B->xgrp, m, ph, 1))) return e;
        e = L->tsize == sizeof(double) ? group_exchange_now_t<double>(B->xgrp, B->stream)
                                       : group_exchange_now_t<float>(B->xgrp, B->stream);
        HIPCHK(hipStreamSynchronize(B->stream));
        if (e) return e;
    }
    return TMK_OK;
}

extern "C" int tmk_batch_add(tmk_batch* B, tmk_level* L, double time, double dt)
{
    if (!B || !L || L->batch || L->amr_stream) return TMK_E_ARG;
    if (!B->members.empty()) {
        tmk_level* L0 = B->members[0];
        if (L0->tsize != L->tsize || L0->ns != L->ns || L0->bx != L->bx || L0->by != L->by || L0->march != L->march)
            return TMK_E_ARG;
    }
    HIPCHK(hipStreamSynchronize(L->ctx->stream));
    int e = batch_detach_all(B);
    if (e) return e;
    L->time = time;
    L->dt = dt;
    L->batch = B;
    B->members.push_back(L);
    B->phase.push_back(L->cur);
    B->stale = true;
    return TMK_OK;
}

extern "C" int tmk_batch_remove(tmk_batch* B, tmk_level* L)
{
    if (!B || !L || L->batch != B) return TMK_E_ARG;
    int e = batch_detach_all(B);
    if (e) return e;
    const auto it = std::find(B->members.begin(), B->members.end(), L);
    const size_t m = it - B->members.begin();
    B->members.erase(it);
    B->phase.erase(B->phase.begin() + m);
    L->batch = nullptr;
    L->own_built = false;  // rebuild its own group at its current parity
    L->own.reset();
    B->stale = true;
    return TMK_OK;
}

extern "C" int tmk_batch_run(tmk_batch* B, int nsteps)
{
    if (!B || nsteps < 0) return TMK_E_ARG;
    int e;
    if (B->stale && (e = batch_rebuild(B))) return e;
    if (!B->built || nsteps == 0) return TMK_OK;
    e = group_run(B->grp, B->stream, nsteps);
    if (e) return e;
    for (tmk_level* L : B->members)
        for (int k = 0; k < nsteps; ++k) L->time = L->time + L->dt;
    return TMK_OK;
}

// ---------------------------------------------------------------------------
// fault stations
// ---------------------------------------------------------------------------
extern "C" int tmk_stations_create(tmk_level* L, int n, const double* x, int capacity, tmk_stations** out)
{
    if (!L || !out || n <= 0 || !x || capacity <= 0 || L->stations || L->batch || L->amr_stream) return TMK_E_ARG;
    *out = nullptr;
    tmk_ctx* c = L->ctx;
    const double dx = c->hp.dx;
    const int nb = c->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
    std::vector<int> patch(n), face(n);
    std::vector<double> d(n);
    for (int k = 0; k < n; ++k) {
        // AMRLevelLinElast.cpp:262-270 (x already includes the domain centre)
        const int i = (int)std::floor(x[k] / dx - 0.5);
        face[k] = i;
        d[k] = -(i + 0.5) * dx + x[k];
        patch[k] = -1;
        for (size_t b = 0; b < L->boxes.size(); ++b)
            if (L->has_b[b] && i >= L->boxes[b].lo[0] && i <= L->boxes[b].hi[0]) patch[k] = (int)b;
        if (patch[k] < 0) return TMK_E_ARG;
    }
    tmk_stations* S = new (std::nothrow) tmk_stations;
    if (!S) return TMK_E_NOMEM;
    int e = TMK_OK;
    const size_t mbytes = (size_t)n * (2 * sizeof(int) + sizeof(double));
    if (!e) e = S->meta.ensure(mbytes);
    if (!e) e = S->rec.ensure((size_t)capacity * n * (1 + nb) * sizeof(double));
    if (!e) e = S->count.ensure(2 * sizeof(int));
    if (!e) {
        char* m = (char*)S->meta.p;
        const int zero[2] = {0, 0};
        hipStream_t st = c->stream;
        if (copy_on(st, m, d.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            copy_on(st, m + n * sizeof(double), patch.data(), n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
            copy_on(st, m + n * (sizeof(double) + sizeof(int)), face.data(), n * sizeof(int),
                    hipMemcpyHostToDevice) != hipSuccess ||
            copy_on(st, S->count.p, zero, sizeof zero, hipMemcpyHostToDevice) != hipSuccess)
            e = TMK_E_HIP;
        S->dev.n = n;
        S->dev.nb = nb;
        S->dev.capacity = capacity;
        S->dev.d = (const double*)m;
        S->dev.patch = (const int*)(m + n * sizeof(double));
        S->dev.face = (const int*)(m + n * (sizeof(double) + sizeof(int)));
        S->dev.rec = (double*)S->rec.p;
        S->dev.count = (int*)S->count.p;
    }
    if (e) {
        S->meta.release(); S->rec.release(); S->count.release();
        delete S;
        return e;
    }
    S->L = L;
    L->stations = S;
    L->own_built = false;  // the next step rebuilds the group (and its graph) with the recorder
    *out = S;
    return TMK_OK;
}

extern "C" int tmk_stations_read(tmk_stations* S, double* records, int max_records, int* nrecords)
{
    if (!S || !S->L || !records || !nrecords || max_records < S->dev.capacity) return TMK_E_ARG;
    hipStream_t st = S->L->ctx->stream;
    int cnt[2] = {0, 0};
    HIPCHK(copy_on(st, cnt, S->count.p, sizeof cnt, hipMemcpyDeviceToHost));
    const int n = std::min(cnt[0], S->dev.capacity);
    if (n > 0)
        HIPCHK(copy_on(st, records, S->rec.p, (size_t)n * S->dev.n * (1 + S->dev.nb) * sizeof(double),
                       hipMemcpyDeviceToHost));
    const int zero[2] = {0, 0};
    HIPCHK(copy_on(st, S->count.p, zero, sizeof zero, hipMemcpyHostToDevice));
    *nrecords = n;
    return cnt[1] ? TMK_E_OVERFLOW : TMK_OK;
}

extern "C" int tmk_stations_destroy(tmk_stations* S)
{
    if (!S) return TMK_OK;
    if (S->L) {
        (void)hipStreamSynchronize(S->L->ctx->stream);
        S->L->stations = nullptr;
        S->L->own_built = false;
    }
    S->meta.release();
    S->rec.release();
    S->count.release();
    delete S;
    return TMK_OK;
}

extern "C" int tmk_batch_member_error(tmk_batch* B, tmk_level* L, int* code)
{
    if (!B || !L || !code || L->batch != B) return TMK_E_ARG;
    HIPCHK(hipStreamSynchronize(B->stream));
    *code = 0;
    if (!B->built) return TMK_OK;
    const size_t m = std::find(B->members.begin(), B->members.end(), L) - B->members.begin();
    if (m >= B->members.size()) return TMK_E_ARG;
    return group_error(B->grp, B->stream, (int)m, code);
}

extern "C" int tmk_batch_done(tmk_batch* B, int* done)
{
    if (!B || !done) return TMK_E_ARG;
    const hipError_t e = hipStreamQuery(B->stream);  // the side stream joins the batch stream
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();
        *done = 0;
        return TMK_OK;
    }
    HIPCHK(e);
    *done = 1;
    return TMK_OK;
}

extern "C" int tmk_batch_sync(tmk_batch* B)
{
    if (!B) return TMK_E_ARG;
    HIPCHK(hipStreamSynchronize(B->stream));
    if (!B->built) return TMK_OK;
    std::vector<Scen> sc(B->members.size());
    HIPCHK(copy_on(B->stream, sc.data(), B->grp.scen_dev, sizeof(Scen) * sc.size(), hipMemcpyDeviceToHost));
    for (const Scen& s : sc)
        if (s.err) return s.err;
    return TMK_OK;
}

// ---------------------------------------------------------------------------
// refinement tagging
// ---------------------------------------------------------------------------
template <typename T>
static int level_tag_t(tmk_level* L, const tmk_tagparams* tp, unsigned char* tags)
{
    tmk_ctx* c = L->ctx;
    hipStream_t st = c->stream;
    const tmk_box dom = c->hp.domain;
    const int nx = dom.hi[0] - dom.lo[0] + 1, ny = dom.hi[1] - dom.lo[1] + 1;
    const size_t n = (size_t)nx * ny;
    int e;
    if ((e = c->bufs[6].ensure(2 * n))) return e;
    unsigned char* raw = (unsigned char*)c->bufs[6].p;
    HIPCHK(hipMemsetAsync(raw, 0, 2 * n, st));
    TagArgs ta;
    ta.refine_thresh = tp->refine_thresh; ta.plastic_thresh = tp->plastic_thresh;
    ta.xc0 = tp->x_coarsen[0]; ta.xc1 = tp->x_coarsen[1]; ta.yc0 = tp->y_coarsen[0]; ta.yc1 = tp->y_coarsen[1];
    ta.slope = tp->slope_coarsen; ta.width = tp->width_coarsen;
    ta.x2c0 = tp->x_coarsen2[0]; ta.x2c1 = tp->x_coarsen2[1]; ta.y2c0 = tp->y_coarsen2[0]; ta.y2c1 = tp->y_coarsen2[1];
    ta.level = tp->level; ta.c2max = tp->coarsen2_max_level;
    ta.stress_thresh = tp->stress_thresh; ta.sliprate_thresh = tp->sliprate_thresh;
    ta.nx = nx; ta.ny = ny;
    // patches_dev[g].in hold the current state
    const Patch<T>* pd = (const Patch<T>*)L->own.patches_dev[L->own.g];
    hipLaunchKernelGGL(k_tag<T>, dim3((unsigned)L->boxes.size(), 64), dim3(256), 0, st, pd, L->own.scen_dev, ta, raw);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_dilate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, raw, raw + n, nx, ny,
                       tp->tag_buffer);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(tags, raw + n, n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return TMK_OK;
}

extern "C" int tmk_level_tag(tmk_level* L, const tmk_tagparams* tp, unsigned char* tags)
{
    if (!L || !tp || !tags || L->batch || L->amr_stream || tp->tag_buffer < 0) return TMK_E_ARG;
    int e = level_prepare(L, L->time, L->dt);
    if (e) return e;
    return L->tsize == sizeof(double) ? level_tag_t<double>(L, tp, tags) : level_tag_t<float>(L, tp, tags);
}

// ---------------------------------------------------------------------------
// hierarchy of levels: the subcycled multi-level step
// (Chombo AMR::timeStep with AMRLevelLinElast::advance / postTimeStep,
//  AMRLevelLinElast.cpp:372-516, and the coarse-fine parts of
//  LinElastLevelGodunov::step, LinElastLevelGodunov.cpp:226-355)
// ---------------------------------------------------------------------------
struct AmrMember;
struct PairLists;

// one level of a hierarchy batch: the members' level-l boxes in one group (shared launches)
struct AmrLevel {
    Group grp;
    std::vector<int> phase;             // member parity at group parity 0
    std::vector<int> poff;              // first descriptor of member m in the group's tables
    // coarse-fine fill of this level from the next coarser one (levels 1 ..)
    DevBuf cells, faces, comps;
    int ncells = 0, nfaces = 0, ncomps = 0;
    // flux register with the next finer level, and the average-down from it (levels .. n-2)
    DevBuf regf, reg, rcells, rscale, avgc, avgf, acomps;
    int nreg = 0, nrcells = 0, navgc = 0, navgf = 0, nacomps = 0;
};

// one scenario of a hierarchy batch: its levels, and the face fluxes and face masks of their
// islands (see Patch::fmask), kept while it is a member
struct AmrMember {
    std::vector<tmk_level*> L;
    std::vector<std::vector<double*>> F[2];          // [direction][level][island]
    std::vector<std::vector<unsigned char*>> fmask;  // [level][island]
    std::vector<PairLists*> lists;                   // [pair l-1, l], member-local indices
    double dt0 = 0.0;
    double tnext = 0.0;                              // coarse clock after the queued steps (tmk_amr_run)
    int err = 0;                                     // first error seen at a detach
    ~AmrMember();
};

// A hierarchy batch: scenarios with the same number of levels and refinement ratio, their
// level-l boxes updated in shared launches, each member with its own parameters, clock and
// dt.  One member is the plain hierarchy of tmk_amr_create.
struct tmk_amr {
    int nlev = 0;
    int r = 2;
    std::vector<AmrMember*> mem;
    std::vector<AmrLevel> lv;
    hipStream_t stream = nullptr;
    bool built = false;   // groups and lists match the members
    hipGraphExec_t gexec = nullptr;  // two coarse steps (every level back to its parity)
    int gpar = 0;                    // level-0 group parity the graph was captured at
    Group xgrp;                      // one-level group for the exchange of an uploaded level
};

// offset of cell (i, j) in the plane of the island array holding box b
static inline int cell_off(const tmk_level* L, int b, int i, int j)
{
    const Island& I = L->isl[L->box_isl[b]];
    return (i - I.abox.lo[0]) + (j - I.abox.lo[1]) * I.unx();
}

// offset of fault face i in the fault-data row of the island holding box b
static inline int face_off(const tmk_level* L, int b, int i) { return i - L->isl[L->box_isl[b]].blo0; }

// descriptor index of the island holding box b (the tiles' descriptor: its flux arrays)
static inline int isl_patch(const tmk_level* L, int b) { return (int)L->boxes.size() + L->box_isl[b]; }

// face offset in the flux array of direction d over bounding box v
static inline int flux_off(const tmk_box& v, int d, int fi, int fj)
{
    const int w = v.hi[0] - v.lo[0] + 1;
    return (fi - v.lo[0]) + (fj - v.lo[1]) * (w + (d == 0));
}

template <typename X>
static int upload_list(DevBuf& buf, const std::vector<X>& v, hipStream_t st)
{
    int e = buf.ensure(std::max<size_t>(sizeof(X) * v.size(), 1));
    if (e) return e;
    HIPCHK(copy_on(st, buf.p, v.data(), sizeof(X) * v.size(), hipMemcpyHostToDevice));
    return TMK_OK;
}

// Coverage maps of the pair (coarse c, fine f).  The 2-D maps cover a window of the coarse
// domain: the coarsened bounding box of the fine boxes grown by kPairMargin cells, which holds
// every coarse cell the pair's lists read (fine ghost cells reach TMK_GHOST / r coarse cells out,
// their interpolation stencils one more, the register faces one).  c_at answers outside the
// window too, from the coarse boxes, so the window's size decides speed and never results.
constexpr int kPairMargin = 4;
struct PairMaps {
    int ncx = 0, ncy = 0;                  // coarse domain
    int wx0 = 0, wy0 = 0, wnx = 0, wny = 0;  // window of own_c / own_f
    const tmk_level* C = nullptr;
    std::vector<int> own_c;   // coarse box holding the cell (window)
    std::vector<int> own_f;   // fine box covering the cell, coarsened (window)
    std::vector<int> own_cb;  // coarse box with fault data holding fault face i (domain)
    std::vector<int> own_fb;  // fine box with fault data covering coarse fault face i (domain)
    bool in_w(int i, int j) const { return i >= wx0 && i < wx0 + wnx && j >= wy0 && j < wy0 + wny; }
    size_t wq(int i, int j) const { return (size_t)(i - wx0) + (size_t)(j - wy0) * wnx; }
    int c_at(int i, int j) const
    {
        if (in_w(i, j)) return own_c[wq(i, j)];
        for (size_t b = 0; b < C->boxes.size(); ++b) {
            const tmk_box& v = C->boxes[b];
            if (i >= v.lo[0] && i <= v.hi[0] && j >= v.lo[1] && j <= v.hi[1]) return (int)b;
        }
        return -1;
    }
    int f_at(int i, int j) const { return in_w(i, j) ? own_f[wq(i, j)] : -1; }  // the window holds every fine box
};

static int pair_maps(const tmk_level* C, const tmk_level* Fl, int r, PairMaps& m)
{
    const tmk_box cd = C->ctx->hp.domain, fd = Fl->ctx->hp.domain;
    if (cd.lo[0] != 0 || cd.lo[1] != 0 || fd.lo[0] != 0 || fd.lo[1] != 0) return TMK_E_NESTING;
    if (fd.hi[0] + 1 != (cd.hi[0] + 1) * r || fd.hi[1] + 1 != (cd.hi[1] + 1) * r) return TMK_E_NESTING;
    if (Fl->ctx->hp.dx * r != C->ctx->hp.dx) return TMK_E_NESTING;
    m.C = C;
    m.ncx = cd.hi[0] + 1;
    m.ncy = cd.hi[1] + 1;
    int x0 = m.ncx, y0 = m.ncy, x1 = -1, y1 = -1;
    for (const tmk_box& v : Fl->boxes) {
        for (int d = 0; d < 2; ++d)
            if (v.lo[d] % r != 0 || (v.hi[d] + 1) % r != 0) return TMK_E_NESTING;
        x0 = std::min(x0, v.lo[0] / r);
        y0 = std::min(y0, v.lo[1] / r);
        x1 = std::max(x1, v.hi[0] / r);
        y1 = std::max(y1, v.hi[1] / r);
    }
    if (x1 < 0) return TMK_E_NESTING;
    m.wx0 = std::max(0, x0 - kPairMargin);
    m.wy0 = std::max(0, y0 - kPairMargin);
    m.wnx = std::min(m.ncx - 1, x1 + kPairMargin) - m.wx0 + 1;
    m.wny = std::min(m.ncy - 1, y1 + kPairMargin) - m.wy0 + 1;
    const size_t n = (size_t)m.wnx * m.wny;
    m.own_c.assign(n, -1);
    m.own_f.assign(n, -1);
    m.own_cb.assign(m.ncx, -1);
    m.own_fb.assign(m.ncx, -1);
    const tmk_box win{{m.wx0, m.wy0}, {m.wx0 + m.wnx - 1, m.wy0 + m.wny - 1}};
    for (size_t b = 0; b < C->boxes.size(); ++b) {
        const tmk_box v = C->boxes[b];
        const tmk_box x = box_and(v, win);
        for (int j = x.lo[1]; j <= x.hi[1]; ++j)
            for (int i = x.lo[0]; i <= x.hi[0]; ++i) m.own_c[m.wq(i, j)] = (int)b;
        if (v.lo[1] == 0)
            for (int i = v.lo[0]; i <= v.hi[0]; ++i) m.own_cb[i] = (int)b;
    }
    for (size_t b = 0; b < Fl->boxes.size(); ++b) {
        const tmk_box v = Fl->boxes[b];
        for (int j = v.lo[1] / r; j <= v.hi[1] / r; ++j)
            for (int i = v.lo[0] / r; i <= v.hi[0] / r; ++i) {
                if (m.own_c[m.wq(i, j)] < 0) return TMK_E_NESTING;
                m.own_f[m.wq(i, j)] = (int)b;
            }
        if (v.lo[1] == 0)
            for (int i = v.lo[0] / r; i <= v.hi[0] / r; ++i) m.own_fb[i] = (int)b;
    }
    return TMK_OK;
}

// register faces of the pair in the order (direction, side, j, i): an uncovered coarse
// cell with the refined region on that side.  fn(d, side, i, j, coarse box, fine box)
template <typename Fn>
static void for_each_reg_face(const PairMaps& m, Fn fn)
{
    const int ncx = m.ncx, ncy = m.ncy;
    for (int d = 0; d < 2; ++d)
        for (int side = 0; side < 2; ++side)
            for (int j = m.wy0; j < m.wy0 + m.wny; ++j)
                for (int i = m.wx0; i < m.wx0 + m.wnx; ++i) {
                    const size_t q = m.wq(i, j);
                    if (m.own_c[q] < 0 || m.own_f[q] >= 0) continue;
                    const int sg = side ? 1 : -1;
                    const int ni = i + (d == 0 ? sg : 0), nj = j + (d == 1 ? sg : 0);
                    if (ni < 0 || ni >= ncx || nj < 0 || nj >= ncy) continue;
                    const int fb = m.f_at(ni, nj);
                    if (fb >= 0) fn(d, side, i, j, m.own_c[q], fb);
                }
}

struct PairLists {
    std::vector<CfCell> cells;
    std::vector<CfFace> faces;
    std::vector<RegFace> regf;
    std::vector<double> rscale;
    std::vector<RefluxCell> rcells;
    std::vector<AvgCell> avgc;
    std::vector<AvgFace> avgf;
};

AmrMember::~AmrMember()
{
    for (PairLists* p : lists) delete p;
}

// the lists of member M's pair (level l-1, level l), with the member's own descriptor indices
static int member_pair_lists(const AmrMember& M, int l, int r, PairLists& P)
{
    const tmk_level* C = M.L[l - 1];
    const tmk_level* Fl = M.L[l];
    PairMaps m;
    int e = pair_maps(C, Fl, r, m);
    if (e) return e;
    const int ncx = m.ncx, ncy = m.ncy;
    const tmk_box fd = Fl->ctx->hp.domain;
    auto in_c = [&](int i, int j) { return i >= 0 && i < ncx && j >= 0 && j < ncy && m.c_at(i, j) >= 0; };
    // coarse-fine fill of the state: one entry per (fine box, ghost cell inside the domain
    // and outside every fine box)
    std::vector<CfCell> cells;
    for (size_t b = 0; b < Fl->boxes.size(); ++b) {
        const tmk_box v = Fl->boxes[b];
        const int i0 = std::max(v.lo[0] - TMK_GHOST, 0), i1 = std::min(v.hi[0] + TMK_GHOST, fd.hi[0]);
        const int j0 = std::max(v.lo[1] - TMK_GHOST, 0), j1 = std::min(v.hi[1] + TMK_GHOST, fd.hi[1]);
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) {
                const int ic = i / r, jc = j / r;
                if (m.f_at(ic, jc) >= 0) continue;  // a cell of the fine level
                CfCell cc;
                memset(&cc, 0, sizeof cc);
                cc.dpatch = (int)b;
                cc.doff = cell_off(Fl, (int)b, i, j);
                for (int jj = -1, k = 0; jj <= 1; ++jj)
                    for (int ii = -1; ii <= 1; ++ii, ++k) {
                        const int ci = ic + ii, cj = jc + jj;
                        cc.cpatch[k] = -1;
                        if (in_c(ci, cj)) {
                            const int cb = m.c_at(ci, cj);
                            cc.cpatch[k] = cb;
                            cc.coff[k] = cell_off(C, cb, ci, cj);
                        }
                    }
                if (cc.cpatch[4] < 0) return TMK_E_NESTING;
                const int lo[2] = {cc.cpatch[3] >= 0, cc.cpatch[1] >= 0};
                const int hi[2] = {cc.cpatch[5] >= 0, cc.cpatch[7] >= 0};
                int flags = 0;
                for (int d = 0; d < 2; ++d) {
                    int kind;
                    if (lo[d] && hi[d]) kind = 0;       // central
                    else if (lo[d]) kind = 1;           // one-sided low
                    else if (hi[d]) kind = 2;           // one-sided high
                    else return TMK_E_NESTING;
                    flags |= kind << (2 * d);
                }
                // the limiter covers the coarse cells off the domain's edge layer
                if (ic >= 1 && ic <= ncx - 2 && jc >= 1 && jc <= ncy - 2) {
                    for (int k = 0; k < 9; ++k)
                        if (cc.cpatch[k] < 0) return TMK_E_NESTING;
                    flags |= 16;
                }
                cc.flags = flags;
                cc.coef0 = -0.5 + ((double)(i - r * ic) + 0.5) / (double)r;
                cc.coef1 = -0.5 + ((double)(j - r * jc) + 0.5) / (double)r;
                cells.push_back(cc);
            }
    }
    std::vector<int> comps;
    for (int k = 0; k < TMK_NCOMP; ++k)
        if (is_stencil(k, Fl->ns)) comps.push_back(k);
    // coarse-fine fill of the fault data
    std::vector<CfFace> faces;
    for (size_t b = 0; b < Fl->boxes.size(); ++b) {
        const tmk_box v = Fl->boxes[b];
        if (!Fl->has_b[b]) continue;
        const int i0 = std::max(v.lo[0] - TMK_GHOST, 0), i1 = std::min(v.hi[0] + TMK_GHOST, fd.hi[0]);
        for (int i = i0; i <= i1; ++i) {
            const int ic = i / r;
            if (m.own_fb[ic] >= 0) continue;
            CfFace ff;
            memset(&ff, 0, sizeof ff);
            ff.dpatch = (int)b;
            ff.doff = face_off(Fl, (int)b, i);
            for (int k = 0; k < 3; ++k) {
                const int ci = ic - 1 + k;
                const int cb = ci >= 0 && ci < ncx ? m.own_cb[ci] : -1;
                ff.cpatch[k] = cb;
                ff.coff[k] = cb >= 0 ? face_off(C, cb, ci) : 0;
            }
            if (ff.cpatch[1] < 0) return TMK_E_NESTING;
            const bool lo = ff.cpatch[0] >= 0, hi = ff.cpatch[2] >= 0;
            if (!lo && !hi) return TMK_E_NESTING;
            ff.kind = lo && hi ? 0 : lo ? 1 : 2;
            ff.coef = -0.5 + ((double)(i - r * ic) + 0.5) / (double)r;
            faces.push_back(ff);
        }
    }
    // flux-register faces, grouped by uncovered coarse cell (a stable sort keeps each
    // cell's faces in the order direction 0 lo, 0 hi, 1 lo, 1 hi)
    struct Ent {
        size_t cell;
        RegFace rf;
        double s;
        int cbox;
    };
    std::vector<Ent> ents;
    const double inv = -1.0 / C->ctx->hp.dx;
    bool bad = false;
    for_each_reg_face(m, [&](int d, int side, int i, int j, int cb, int fb) {
        Ent en;
        memset(&en, 0, sizeof en);
        en.cell = (size_t)j * ncx + i;
        const int fi = d == 0 ? i + side : i, fj = d == 1 ? j + side : j;
        en.rf.cpatch = isl_patch(C, cb);
        en.rf.cface = flux_off(C->isl[C->box_isl[cb]].bbox, d, fi, fj);
        en.rf.d = d;
        en.rf.fpatch = isl_patch(Fl, fb);
        for (int mm = 0; mm < r; ++mm) {
            const int ffi = d == 0 ? fi * r : fi * r + mm, ffj = d == 0 ? fj * r + mm : fj * r;
            en.rf.fface[mm] = flux_off(Fl->isl[Fl->box_isl[fb]].bbox, d, ffi, ffj);
        }
        if (!M.F[d][l - 1][C->box_isl[cb]] || !M.F[d][l][Fl->box_isl[fb]]) bad = true;
        en.cbox = cb;
        en.s = inv * (side ? 1.0 : -1.0);
        ents.push_back(en);
    });
    if (bad) return TMK_E_NESTING;
    std::stable_sort(ents.begin(), ents.end(), [](const Ent& a, const Ent& b) { return a.cell < b.cell; });
    std::vector<RegFace> regf(ents.size());
    std::vector<double> rscale(ents.size());
    std::vector<RefluxCell> rcells;
    for (size_t k = 0; k < ents.size(); ++k) {
        regf[k] = ents[k].rf;
        rscale[k] = ents[k].s;
        if (k == 0 || ents[k].cell != ents[k - 1].cell) {
            RefluxCell rc;
            const int i = (int)(ents[k].cell % ncx), j = (int)(ents[k].cell / ncx);
            rc.cpatch = ents[k].cbox;
            rc.coff = cell_off(C, rc.cpatch, i, j);
            rc.e0 = (int)k;
            rc.e1 = (int)k + 1;
            rcells.push_back(rc);
        } else {
            rcells.back().e1 = (int)k + 1;
        }
    }
    // average-down of the state and of the fault data
    std::vector<AvgCell> avgc;
    std::vector<AvgFace> avgf;
    for (size_t b = 0; b < Fl->boxes.size(); ++b) {
        const tmk_box v = Fl->boxes[b];
        for (int jc = v.lo[1] / r; jc <= v.hi[1] / r; ++jc)
            for (int ic = v.lo[0] / r; ic <= v.hi[0] / r; ++ic) {
                AvgCell a;
                a.cpatch = m.c_at(ic, jc);
                a.coff = cell_off(C, a.cpatch, ic, jc);
                a.fpatch = (int)b;
                a.foff = cell_off(Fl, (int)b, ic * r, jc * r);
                avgc.push_back(a);
            }
        if (!Fl->has_b[b]) continue;
        for (int ic = v.lo[0] / r; ic <= v.hi[0] / r; ++ic) {
            AvgFace a;
            a.cpatch = m.own_cb[ic];
            if (a.cpatch < 0) return TMK_E_NESTING;
            a.coff = face_off(C, a.cpatch, ic);
            a.fpatch = (int)b;
            a.foff = face_off(Fl, (int)b, ic * r);
            avgf.push_back(a);
        }
    }
    P.cells.swap(cells);
    P.faces.swap(faces);
    P.regf.swap(regf);
    P.rscale.swap(rscale);
    P.rcells.swap(rcells);
    P.avgc.swap(avgc);
    P.avgf.swap(avgf);
    return TMK_OK;
}

// face fluxes and masks of member M on its levels' descriptors (both parities)
template <typename T>
static void member_set_flux_ptrs(AmrMember& M, bool on)
{
    for (size_t l = 0; l < M.L.size(); ++l) {
        tmk_level* L = M.L[l];
        const int nbox = (int)L->boxes.size();
        for (int par = 0; par < 2; ++par) {
            Patch<T>* pts = (Patch<T>*)L->hpatch[par].data();
            for (int q = 0; q < L->npatch; ++q) {
                const int k = q - nbox;
                const bool isl = on && k >= 0 && k < (int)M.F[0][l].size();
                pts[q].F[0] = isl ? M.F[0][l][k] : nullptr;
                pts[q].F[1] = isl ? M.F[1][l][k] : nullptr;
                pts[q].fmask = isl ? M.fmask[l][k] : nullptr;
            }
        }
        L->own_built = false;
    }
}

static void member_release(AmrMember& M)
{
    for (int d = 0; d < 2; ++d)
        for (auto& v : M.F[d])
            for (double* p : v) dev_free(p);
    for (auto& v : M.fmask)
        for (unsigned char* p : v) dev_free(p);
    for (int d = 0; d < 2; ++d) M.F[d].assign(M.L.size(), {});
    M.fmask.assign(M.L.size(), {});
    if (!M.L.empty()) {
        if (M.L[0]->tsize == sizeof(double)) member_set_flux_ptrs<double>(M, false);
        else member_set_flux_ptrs<float>(M, false);
    }
    for (tmk_level* L : M.L) L->amr_stream = nullptr;
}

// flux arrays and face masks of every island of the member's levels (mask bit 0: a cell of
// the level, bit 1: under the next finer level), over the island's bounding box grown by one
static int member_layouts(AmrMember& M, int r)
{
    const int n = (int)M.L.size();
    for (int d = 0; d < 2; ++d) M.F[d].assign(n, {});
    M.fmask.assign(n, {});
    for (int l = 0; l < n; ++l) {
        const tmk_level* L = M.L[l];
        const size_t ni = L->isl.size();
        M.F[0][l].assign(ni, nullptr);
        M.F[1][l].assign(ni, nullptr);
        M.fmask[l].assign(ni, nullptr);
        std::vector<tmk_box> finer;
        if (l + 1 < n)
            for (const tmk_box& b : M.L[l + 1]->boxes)
                finer.push_back(tmk_box{{b.lo[0] / r, b.lo[1] / r}, {b.hi[0] / r, b.hi[1] / r}});
        for (size_t k = 0; k < ni; ++k) {
            const Island& I = L->isl[k];
            const size_t w = I.bbox.hi[0] - I.bbox.lo[0] + 1, h = I.bbox.hi[1] - I.bbox.lo[1] + 1;
            for (int d = 0; d < 2; ++d) {
                void* p = nullptr;
                const size_t nf = d == 0 ? (w + 1) * h : w * (h + 1);
                if (dev_alloc(&p, sizeof(double) * TMK_NCOMP * nf) != hipSuccess) return TMK_E_NOMEM;
                M.F[d][l][k] = (double*)p;
            }
            const size_t mw = w + 2, mh = h + 2;
            std::vector<unsigned char> mk(mw * mh, 0);
            auto mark = [&](const tmk_box& b, unsigned char bit) {
                const int i0 = std::max(b.lo[0], I.bbox.lo[0] - 1), i1 = std::min(b.hi[0], I.bbox.hi[0] + 1);
                const int j0 = std::max(b.lo[1], I.bbox.lo[1] - 1), j1 = std::min(b.hi[1], I.bbox.hi[1] + 1);
                for (int j = j0; j <= j1; ++j)
                    for (int i = i0; i <= i1; ++i)
                        mk[(size_t)(i - I.bbox.lo[0] + 1) + (size_t)(j - I.bbox.lo[1] + 1) * mw] |= bit;
            };
            for (const tmk_box& b : L->boxes) mark(b, 1);
            for (const tmk_box& b : finer) mark(b, 2);
            void* p = nullptr;
            if (dev_alloc(&p, mk.size()) != hipSuccess) return TMK_E_NOMEM;
            M.fmask[l][k] = (unsigned char*)p;
            HIPCHK(copy_on(L->ctx->stream, p, mk.data(), mk.size(), hipMemcpyHostToDevice));
        }
    }
    return TMK_OK;
}

// pull the members' clocks, parities and errors out of the groups before the layout changes
static int amr_detach(tmk_amr* A)
{
    HIPCHK(hipStreamSynchronize(A->stream));
    if (A->gexec) (void)hipGraphExecDestroy(A->gexec);
    A->gexec = nullptr;
    if (!A->built) return TMK_OK;
    for (int l = 0; l < A->nlev; ++l) {
        AmrLevel& a = A->lv[l];
        std::vector<Scen> sc(A->mem.size());
        if (!sc.empty())
            HIPCHK(copy_on(A->stream, sc.data(), a.grp.scen_dev, sizeof(Scen) * sc.size(), hipMemcpyDeviceToHost));
        for (size_t m = 0; m < A->mem.size(); ++m) {
            tmk_level* L = A->mem[m]->L[l];
            L->cur = a.phase[m] ^ a.grp.g;
            L->time = sc[m].time;
            if (sc[m].err && !A->mem[m]->err) A->mem[m]->err = sc[m].err;
        }
        a.grp.reset();
    }
    A->built = false;
    return TMK_OK;
}

// groups and lists for the current members (lazily, by the first run after a change)
template <typename T>
static int amr_build_t(tmk_amr* A)
{
    const int nm = (int)A->mem.size();
    int e;
    // levels uploaded since their last step: exchange their ghosts first
    for (AmrMember* M : A->mem)
        for (tmk_level* L : M->L) {
            if (!L->dirty) continue;
            L->dirty = false;
            std::vector<tmk_level*> one{L};
            std::vector<int> ph{L->cur};
            if ((e = group_build(A->xgrp, one, ph, 1))) return e;
            if ((e = group_exchange_now_t<T>(A->xgrp, A->stream))) return e;
            HIPCHK(hipStreamSynchronize(A->stream));
        }
    for (int l = 0; l < A->nlev; ++l) {
        AmrLevel& a = A->lv[l];
        std::vector<tmk_level*> ls;
        a.phase.assign(nm, 0);
        a.poff.assign(nm, 0);
        int off = 0;
        for (int m = 0; m < nm; ++m) {
            tmk_level* L = A->mem[m]->L[l];
            ls.push_back(L);
            a.phase[m] = L->cur;
            a.poff[m] = off;
            off += L->npatch;
        }
        if ((e = group_build(a.grp, ls, a.phase, 1))) return e;
    }
    for (int l = 1; l < A->nlev; ++l) {
        AmrLevel& c = A->lv[l - 1];
        AmrLevel& f = A->lv[l];
        PairLists all;
        for (int m = 0; m < nm; ++m) {
            const PairLists& P = *A->mem[m]->lists[l];  // built when the member joined
            const int pc = c.poff[m], pf = f.poff[m];
            for (CfCell x : P.cells) {
                x.dpatch += pf;
                for (int k = 0; k < 9; ++k)
                    if (x.cpatch[k] >= 0) x.cpatch[k] += pc;
                x.m = m;
                all.cells.push_back(x);
            }
            for (CfFace x : P.faces) {
                x.dpatch += pf;
                for (int k = 0; k < 3; ++k)
                    if (x.cpatch[k] >= 0) x.cpatch[k] += pc;
                x.m = m;
                all.faces.push_back(x);
            }
            const int e0 = (int)all.regf.size();
            for (RegFace x : P.regf) {
                x.cpatch += pc;
                x.fpatch += pf;
                x.m = m;
                all.regf.push_back(x);
            }
            all.rscale.insert(all.rscale.end(), P.rscale.begin(), P.rscale.end());
            for (RefluxCell x : P.rcells) {
                x.cpatch += pc;
                x.e0 += e0;
                x.e1 += e0;
                all.rcells.push_back(x);
            }
            for (AvgCell x : P.avgc) {
                x.cpatch += pc;
                x.fpatch += pf;
                all.avgc.push_back(x);
            }
            for (AvgFace x : P.avgf) {
                x.cpatch += pc;
                x.fpatch += pf;
                all.avgf.push_back(x);
            }
        }
        std::vector<int> comps, acomps;
        for (int k = 0; k < TMK_NCOMP; ++k)
            if (is_stencil(k, A->mem[0]->L[l]->ns)) comps.push_back(k);
        // the average-down skips the anti-plane planes of mode II levels: zero on every level
        // (uploads and fills refuse other values there, the update never writes them, and the
        // members of a batch agree on the mode), so their average is the zero already there
        for (int k = 0; k < TMK_NCOMP; ++k)
            if (A->mem[0]->L[l]->ns == 3 || !(k % 11 == 2 || k % 11 == 7 || k % 11 == 8)) acomps.push_back(k);
        c.nacomps = (int)acomps.size();
        f.ncells = (int)all.cells.size();
        f.nfaces = (int)all.faces.size();
        f.ncomps = (int)comps.size();
        c.nreg = (int)all.regf.size();
        c.nrcells = (int)all.rcells.size();
        c.navgc = (int)all.avgc.size();
        c.navgf = (int)all.avgf.size();
        hipStream_t st = A->stream;
        if ((e = upload_list(f.cells, all.cells, st)) || (e = upload_list(f.faces, all.faces, st)) ||
            (e = upload_list(f.comps, comps, st)) || (e = upload_list(c.regf, all.regf, st)) ||
            (e = upload_list(c.rcells, all.rcells, st)) || (e = upload_list(c.rscale, all.rscale, st)) ||
            (e = upload_list(c.avgc, all.avgc, st)) || (e = upload_list(c.avgf, all.avgf, st)) ||
            (e = upload_list(c.acomps, acomps, st)))
            return e;
        if ((e = c.reg.ensure(std::max<size_t>(sizeof(double) * TMK_NCOMP * all.regf.size(), 1)))) return e;
    }
    A->built = nm > 0;
    return TMK_OK;
}

static int amr_build(tmk_amr* A)
{
    if (A->mem.empty()) return TMK_OK;
    return A->mem[0]->L[0]->tsize == sizeof(double) ? amr_build_t<double>(A) : amr_build_t<float>(A);
}

extern "C" int tmk_amr_batch_create(int nlevels, int ref_ratio, tmk_amr** out)
{
    if (!out || nlevels < 1 || ref_ratio < 2 || ref_ratio > TMK_AMR_MAXR) return TMK_E_ARG;
    *out = nullptr;
    tmk_amr* A = new (std::nothrow) tmk_amr;
    if (!A) return TMK_E_NOMEM;
    A->nlev = nlevels;
    A->r = ref_ratio;
    A->lv.resize(nlevels);
    if (hipStreamCreateWithFlags(&A->stream, hipStreamNonBlocking) != hipSuccess) {
        delete A;
        return TMK_E_HIP;
    }
    *out = A;
    return TMK_OK;
}

// join: one scenario's levels (coarsest first), its clock and coarse step
extern "C" int tmk_amr_add(tmk_amr* A, tmk_level* const* levels, double time, double dt0)
{
    if (!A || !levels || !(dt0 > 0.0)) return TMK_E_ARG;
    for (int l = 0; l < A->nlev; ++l) {
        tmk_level* L = levels[l];
        if (!L || L->batch || L->amr_stream || L->stations) return TMK_E_ARG;
        if (L->tsize != levels[0]->tsize) return TMK_E_ARG;
        for (int k = 0; k < l; ++k)
            if (levels[k] == L) return TMK_E_ARG;
    }
    if (!A->mem.empty()) {
        // level l of every member runs in one group: the same arithmetic and update shape
        for (int l = 0; l < A->nlev; ++l) {
            const tmk_level* a = A->mem[0]->L[l];
            const tmk_level* b = levels[l];
            if (a->tsize != b->tsize || a->ns != b->ns || a->bx != b->bx || a->by != b->by || a->march != b->march)
                return TMK_E_ARG;
        }
    }
    {   // level 0 covers its domain
        const tmk_level* L0 = levels[0];
        const tmk_box d = L0->ctx->hp.domain;
        long long cells = 0;
        for (const tmk_box& b : L0->boxes) cells += (long long)(b.hi[0] - b.lo[0] + 1) * (b.hi[1] - b.lo[1] + 1);
        if (cells != (long long)(d.hi[0] - d.lo[0] + 1) * (d.hi[1] - d.lo[1] + 1)) return TMK_E_NESTING;
    }
    for (int l = 1; l < A->nlev; ++l) {
        PairMaps m;
        int e = pair_maps(levels[l - 1], levels[l], A->r, m);
        if (e) return e;
    }
    int e = amr_detach(A);
    if (e) return e;
    AmrMember* M = new (std::nothrow) AmrMember;
    if (!M) return TMK_E_NOMEM;
    M->L.assign(levels, levels + A->nlev);
    M->dt0 = dt0;
    for (tmk_level* L : M->L) HIPCHK(hipStreamSynchronize(L->ctx->stream));
    e = member_layouts(*M, A->r);
    // the member's lists (which also check nesting), kept while it is a member
    M->lists.assign(A->nlev, nullptr);
    for (int l = 1; l < A->nlev && !e; ++l) {
        M->lists[l] = new (std::nothrow) PairLists;
        e = M->lists[l] ? member_pair_lists(*M, l, A->r, *M->lists[l]) : TMK_E_NOMEM;
    }
    if (e) {
        member_release(*M);
        delete M;
        return e;
    }
    if (M->L[0]->tsize == sizeof(double)) member_set_flux_ptrs<double>(*M, true);
    else member_set_flux_ptrs<float>(*M, true);
    double dt = dt0;
    for (tmk_level* L : M->L) {
        L->amr_stream = A->stream;
        L->time = time;
        L->dt = dt;
        dt = dt / A->r;
    }
    A->mem.push_back(M);
    return TMK_OK;
}

static int amr_find(tmk_amr* A, const tmk_level* L0)
{
    for (size_t m = 0; m < A->mem.size(); ++m)
        if (A->mem[m]->L[0] == L0) return (int)m;
    return -1;
}

// leave: the scenario's levels are plain levels again
extern "C" int tmk_amr_remove(tmk_amr* A, tmk_level* level0)
{
    if (!A || !level0) return TMK_E_ARG;
    const int m = amr_find(A, level0);
    if (m < 0) return TMK_E_ARG;
    int e = amr_detach(A);
    if (e) return e;
    AmrMember* M = A->mem[m];
    A->mem.erase(A->mem.begin() + m);
    member_release(*M);
    delete M;
    return TMK_OK;
}

extern "C" int tmk_amr_destroy(tmk_amr* A)
{
    if (!A) return TMK_OK;
    (void)amr_detach(A);
    for (AmrMember* M : A->mem) {
        member_release(*M);
        delete M;
    }
    for (AmrLevel& a : A->lv) a.grp.release();
    A->xgrp.release();
    if (A->stream) (void)hipStreamDestroy(A->stream);
    delete A;
    return TMK_OK;
}

extern "C" int tmk_amr_create(int nlevels, tmk_level* const* levels, int ref_ratio, tmk_amr** out)
{
    if (!out || !levels) return TMK_E_ARG;
    int e = tmk_amr_batch_create(nlevels, ref_ratio, out);
    if (e) return e;
    // the clock is set by the first tmk_amr_run
    if ((e = tmk_amr_add(*out, levels, levels[0]->time, 1.0))) {
        tmk_amr_destroy(*out);
        *out = nullptr;
        return e;
    }
    (*out)->mem[0]->dt0 = 0.0;
    return TMK_OK;
}

static inline dim3 amr_grid(long long n) { return dim3((unsigned)((n + 255) / 256)); }

// AMRLevelLinElast::advance of level l for every member: coarse-fine fill, the level step
// (update with flux capture, exchange, clock), register increments
template <typename T>
static int amr_advance_t(tmk_amr* A, int l)
{
    AmrLevel& a = A->lv[l];
    Group& G = a.grp;
    hipStream_t st = A->stream;
    if (l > 0) {
        Group& GC = A->lv[l - 1].grp;
        const Patch<T>* fine = (const Patch<T>*)G.patches_dev[G.g];
        const Patch<T>* cnew = (const Patch<T>*)GC.patches_dev[GC.g];      // state after the coarse step
        const Patch<T>* cold = (const Patch<T>*)GC.patches_dev[GC.g ^ 1];  // its input: the state before
        if (a.ncells > 0) {
            hipLaunchKernelGGL(k_amr_fill_state<T>, amr_grid((long long)a.ncells * a.ncomps), dim3(256), 0, st,
                               (const CfCell*)a.cells.p, a.ncells, (const int*)a.comps.p, a.ncomps, fine, cold, cnew,
                               G.scen_dev, GC.scen_dev, A->r, -1.0);
            HIPCHK(hipGetLastError());
        }
        if (a.nfaces > 0) {
            hipLaunchKernelGGL(k_amr_fill_fault<T>, amr_grid((long long)a.nfaces * TMK_NBDRY_RS), dim3(256), 0, st,
                               (const CfFace*)a.faces.p, a.nfaces, TMK_NBDRY_RS, fine, cold, cnew, G.scen_dev,
                               GC.scen_dev, A->r, -1.0);
            HIPCHK(hipGetLastError());
        }
    }
    const Patch<T>* pts = (const Patch<T>*)G.patches_dev[G.g];  // flux pointers are the same at both parities
    int e = group_enqueue_step(G, st);
    if (e) return e;
    for (AmrMember* M : A->mem) {
        tmk_level* L = M->L[l];
        L->cur ^= 1;
        L->time = L->time + L->dt;
    }
    if (l + 1 < A->nlev && a.nreg > 0) {
        hipLaunchKernelGGL(k_amr_reg_coarse<T>, amr_grid((long long)a.nreg * TMK_NCOMP), dim3(256), 0, st,
                           (const RegFace*)a.regf.p, a.nreg, (double*)a.reg.p, pts, (const Scen*)G.scen_dev);
        HIPCHK(hipGetLastError());
    }
    if (l > 0) {
        AmrLevel& c = A->lv[l - 1];
        if (c.nreg > 0) {
            hipLaunchKernelGGL(k_amr_reg_fine<T>, amr_grid((long long)c.nreg * TMK_NCOMP), dim3(256), 0, st,
                               (const RegFace*)c.regf.p, c.nreg, (double*)c.reg.p, pts, A->r, (const Scen*)G.scen_dev);
            HIPCHK(hipGetLastError());
        }
    }
    return TMK_OK;
}

// AMRLevelLinElast::postTimeStep of level l: reflux, average-down of the state and the
// fault data from level l+1, then the same-level exchange of the changed state
template <typename T>
static int amr_post_t(tmk_amr* A, int l)
{
    if (l + 1 >= A->nlev) return TMK_OK;
    AmrLevel& a = A->lv[l];
    Group& G = a.grp;
    Group& GF = A->lv[l + 1].grp;
    hipStream_t st = A->stream;
    const Patch<T>* cur = (const Patch<T>*)G.patches_dev[G.g];
    const Patch<T>* fcur = (const Patch<T>*)GF.patches_dev[GF.g];
    const int r = A->r;
    if (a.nrcells > 0) {
        hipLaunchKernelGGL(k_amr_reflux<T>, amr_grid((long long)a.nrcells * TMK_NCOMP), dim3(256), 0, st,
                           (const RefluxCell*)a.rcells.p, a.nrcells, (const double*)a.reg.p,
                           (const double*)a.rscale.p, a.nreg, cur);
        HIPCHK(hipGetLastError());
    }
    if (a.navgc > 0) {
        hipLaunchKernelGGL(k_amr_avg_state<T>, amr_grid((long long)a.navgc * a.nacomps), dim3(256), 0, st,
                           (const AvgCell*)a.avgc.p, a.navgc, (const int*)a.acomps.p, a.nacomps, cur, fcur, r,
                           1.0 / (r * r));
        HIPCHK(hipGetLastError());
    }
    if (a.navgf > 0) {
        hipLaunchKernelGGL(k_amr_avg_fault<T>, amr_grid((long long)a.navgf * TMK_NBDRY_RS), dim3(256), 0, st,
                           (const AvgFace*)a.avgf.p, a.navgf, TMK_NBDRY_RS, cur, fcur, r, 1.0 / r);
        HIPCHK(hipGetLastError());
    }
    return group_exchange_now_t<T>(G, st);
}

template <typename T>
static int amr_time_step_t(tmk_amr* A, int l)
{
    int e = amr_advance_t<T>(A, l);
    if (e) return e;
    if (l + 1 < A->nlev)
        for (int k = 0; k < A->r; ++k)
            if ((e = amr_time_step_t<T>(A, l + 1))) return e;
    return amr_post_t<T>(A, l);
}

static int amr_time_step(tmk_amr* A)
{
    return A->mem[0]->L[0]->tsize == sizeof(double) ? amr_time_step_t<double>(A, 0) : amr_time_step_t<float>(A, 0);
}

// nsteps coarse steps of every member, each from its own clock with its own dt
extern "C" int tmk_amr_advance(tmk_amr* A, int nsteps)
{
    if (!A || nsteps < 0) return TMK_E_ARG;
    if (A->mem.empty() || nsteps == 0) return TMK_OK;
    int e;
    if (!A->built && (e = amr_build(A))) return e;
    Group& G0 = A->lv[0].grp;
    while (nsteps > 0) {
        if (nsteps >= 2 && !A->gexec) {
            // capture two coarse steps: every level returns to its parity
            std::vector<std::vector<double>> t0;
            for (AmrMember* M : A->mem) {
                t0.emplace_back();
                for (tmk_level* L : M->L) t0.back().push_back(L->time);
            }
            // buffer parities of every level group and member level: nothing runs during the
            // capture, so they are restored after it, also when a captured step fails
            std::vector<int> gpar0, cur0;
            for (AmrLevel& lv : A->lv) gpar0.push_back(lv.grp.g);
            for (AmrMember* M : A->mem)
                for (tmk_level* L : M->L) cur0.push_back(L->cur);
            hipGraph_t gr;
            HIPCHK(hipStreamBeginCapture(A->stream, hipStreamCaptureModeThreadLocal));
            A->gpar = G0.g;
            e = amr_time_step(A);
            if (!e) e = amr_time_step(A);
            hipError_t ce = hipStreamEndCapture(A->stream, &gr);
            for (size_t m = 0; m < A->mem.size(); ++m)  // nothing ran yet
                for (size_t l = 0; l < A->mem[m]->L.size(); ++l) A->mem[m]->L[l]->time = t0[m][l];
            for (size_t l = 0; l < A->lv.size(); ++l) A->lv[l].grp.g = gpar0[l];
            {
                size_t q = 0;
                for (AmrMember* M : A->mem)
                    for (tmk_level* L : M->L) L->cur = cur0[q++];
            }
            if (e) return e;
            if (ce != hipSuccess) return TMK_E_HIP;
            hipError_t ie = hipGraphInstantiate(&A->gexec, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            if (ie != hipSuccess) return TMK_E_HIP;
        }
        if (nsteps >= 2 && G0.g == A->gpar) {
            HIPCHK(hipGraphLaunch(A->gexec, A->stream));
            for (AmrMember* M : A->mem)
                for (size_t l = 0; l < M->L.size(); ++l) {
                    tmk_level* L = M->L[l];
                    long long k = 2;
                    for (size_t q = 0; q < l; ++q) k *= A->r;
                    for (long long s = 0; s < k; ++s) L->time = L->time + L->dt;
                }
            nsteps -= 2;
            continue;
        }
        if ((e = amr_time_step(A))) return e;
        --nsteps;
    }
    return TMK_OK;
}

// one-scenario hierarchy: nsteps coarse steps of dt0; the level clocks are (re)set to time
// when it differs from the hierarchy's current time or dt0 changes
extern "C" int tmk_amr_run(tmk_amr* A, int nsteps, double time, double dt0)
{
    if (!A || nsteps < 0 || !(dt0 > 0.0) || A->mem.size() != 1) return TMK_E_ARG;
    AmrMember* M = A->mem[0];
    const bool restart = M->dt0 == 0.0 || time != M->tnext;
    if (restart || dt0 != M->dt0) {
        int e = amr_detach(A);  // reads the device clocks back
        if (e) return e;
        // another time restarts every level's clock there; a new dt at the current time keeps
        // them running, each level's clock the sum of its own steps (AMRLevel's m_time)
        double dt = dt0;
        for (tmk_level* L : M->L) {
            if (restart) L->time = time;
            L->dt = dt;
            dt = dt / A->r;
        }
        M->dt0 = dt0;
        if (restart) M->tnext = time;
    }
    const int e = tmk_amr_advance(A, nsteps);
    // the coarse clock the device reaches: k_advance_time's additions, in the same order
    for (int k = 0; k < nsteps; ++k) M->tnext = M->tnext + dt0;
    return e;
}

extern "C" int tmk_amr_run_timed(tmk_amr* A, int nsteps, double time, double dt0, double* ms)
{
    if (!A || !ms) return TMK_E_ARG;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    int e = TMK_OK;
    // the first run after a change builds the groups (host work) before any device work
    if (A->mem.size() == 1 && (A->mem[0]->dt0 == 0.0 || time != A->mem[0]->tnext || dt0 != A->mem[0]->dt0))
        e = tmk_amr_run(A, 0, time, dt0);
    if (!e && !A->built) e = amr_build(A);
    if (!e && hipEventRecord(a, A->stream) != hipSuccess) e = TMK_E_HIP;
    if (!e) e = tmk_amr_run(A, nsteps, time, dt0);
    if (!e && hipEventRecord(b, A->stream) != hipSuccess) e = TMK_E_HIP;
    if (!e && hipEventSynchronize(b) != hipSuccess) e = TMK_E_HIP;
    float f = 0.f;
    if (!e && hipEventElapsedTime(&f, a, b) != hipSuccess) e = TMK_E_HIP;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!e) *ms = f;
    return e;
}

// waits; *code = the first error member level0 reported (0 = none)
extern "C" int tmk_amr_member_error(tmk_amr* A, tmk_level* level0, int* code)
{
    if (!A || !level0 || !code) return TMK_E_ARG;
    const int m = amr_find(A, level0);
    if (m < 0) return TMK_E_ARG;
    HIPCHK(hipStreamSynchronize(A->stream));
    *code = A->mem[m]->err;
    if (*code || !A->built) return TMK_OK;
    for (AmrLevel& a : A->lv) {
        int c = 0;
        int e = group_error(a.grp, A->stream, m, &c);
        if (e) return e;
        if (c) {
            *code = c;
            return TMK_OK;
        }
    }
    return TMK_OK;
}

// waits for the hierarchy's work; returns the first error any member reported.  As in a
// batch, a member's error stays with it (tmk_amr_member_error) until it leaves.
extern "C" int tmk_amr_done(tmk_amr* A, int* done)
{
    if (!A || !done) return TMK_E_ARG;
    const hipError_t e = hipStreamQuery(A->stream);
    if (e == hipErrorNotReady) {
        (void)hipGetLastError();
        *done = 0;
        return TMK_OK;
    }
    HIPCHK(e);
    *done = 1;
    return TMK_OK;
}

extern "C" int tmk_amr_sync(tmk_amr* A)
{
    if (!A) return TMK_E_ARG;
    HIPCHK(hipStreamSynchronize(A->stream));
    int first = 0;
    for (AmrMember* M : A->mem)
        if (M->err && !first) first = M->err;
    if (!A->built || first) return first;
    for (AmrLevel& a : A->lv) {
        std::vector<Scen> sc(A->mem.size());
        HIPCHK(copy_on(A->stream, sc.data(), a.grp.scen_dev, sizeof(Scen) * sc.size(), hipMemcpyDeviceToHost));
        for (size_t m = 0; m < sc.size() && !first; ++m)
            if (sc[m].err) first = sc[m].err;
    }
    return first;
}

// ---------------------------------------------------------------------------
// regrid: boxes of the next finer level from tags, and the data of a new level
// ---------------------------------------------------------------------------

// Berger-Rigoutsos clustering (the method of Chombo's BRMeshRefine, absent here): the
// tags' bounding box is accepted when the tags fill at least fill_ratio of it; otherwise
// it is cut at the hole of the signatures nearest its middle, else at the strongest sign
// change of the signatures' second difference, else in half along its longer side.
struct TagField {
    int nx = 0, ny = 0;
    std::vector<unsigned char> t;
    bool at(int i, int j) const { return t[(size_t)j * nx + i] != 0; }
};

static void br_cluster(const TagField& g, tmk_box b, double fill, std::vector<tmk_box>& out)
{
    long long cnt = 0;
    tmk_box tb{{b.hi[0] + 1, b.hi[1] + 1}, {b.lo[0] - 1, b.lo[1] - 1}};
    for (int j = b.lo[1]; j <= b.hi[1]; ++j)
        for (int i = b.lo[0]; i <= b.hi[0]; ++i)
            if (g.at(i, j)) {
                ++cnt;
                tb.lo[0] = std::min(tb.lo[0], i); tb.lo[1] = std::min(tb.lo[1], j);
                tb.hi[0] = std::max(tb.hi[0], i); tb.hi[1] = std::max(tb.hi[1], j);
            }
    if (cnt == 0) return;
    b = tb;
    const int w = b.hi[0] - b.lo[0] + 1, h = b.hi[1] - b.lo[1] + 1;
    if ((double)cnt >= fill * (double)w * h) {
        out.push_back(b);
        return;
    }
    std::vector<long long> sig[2] = {std::vector<long long>(w, 0), std::vector<long long>(h, 0)};
    for (int j = b.lo[1]; j <= b.hi[1]; ++j)
        for (int i = b.lo[0]; i <= b.hi[0]; ++i)
            if (g.at(i, j)) { ++sig[0][i - b.lo[0]]; ++sig[1][j - b.lo[1]]; }
    // a cut at k in direction d splits [lo, lo + k - 1] and [lo + k, hi]
    int best_d = -1, best_k = 0;
    double best_dist = 1e300;
    for (int d = 0; d < 2; ++d) {
        const int n = d == 0 ? w : h;
        for (int k = 1; k < n - 1; ++k)
            if (sig[d][k] == 0) {
                const double dist = fabs((k + 0.5) - 0.5 * n);
                if (dist < best_dist) { best_dist = dist; best_d = d; best_k = k; }
            }
    }
    if (best_d >= 0) {
        // split around the hole: the hole row itself holds no tags
        tmk_box a = b, c = b;
        a.hi[best_d] = b.lo[best_d] + best_k - 1;
        c.lo[best_d] = b.lo[best_d] + best_k + 1;
        br_cluster(g, a, fill, out);
        br_cluster(g, c, fill, out);
        return;
    }
    long long best_jump = 0;
    best_dist = 1e300;
    for (int d = 0; d < 2; ++d) {
        const int n = d == 0 ? w : h;
        if (n < 4) continue;
        std::vector<long long> lap(n, 0);
        for (int k = 1; k < n - 1; ++k) lap[k] = sig[d][k - 1] - 2 * sig[d][k] + sig[d][k + 1];
        for (int k = 1; k < n - 2; ++k) {
            if (!((lap[k] < 0 && lap[k + 1] > 0) || (lap[k] > 0 && lap[k + 1] < 0))) continue;
            const long long jump = std::llabs(lap[k + 1] - lap[k]);
            const double dist = fabs((k + 1) - 0.5 * n);
            if (jump > best_jump || (jump == best_jump && dist < best_dist)) {
                best_jump = jump; best_dist = dist; best_d = d; best_k = k + 1;
            }
        }
    }
    if (best_d < 0) {
        best_d = w >= h ? 0 : 1;
        best_k = (best_d == 0 ? w : h) / 2;
    }
    tmk_box a = b, c = b;
    a.hi[best_d] = b.lo[best_d] + best_k - 1;
    c.lo[best_d] = b.lo[best_d] + best_k;
    br_cluster(g, a, fill, out);
    br_cluster(g, c, fill, out);
}

extern "C" int tmk_make_boxes(int nx, int ny, const unsigned char* tags, int ncoarse, const tmk_box* coarse_boxes,
                              const tmk_meshparams* mp, int max_out, tmk_box* out, int* nout)
{
    if (nx <= 0 || ny <= 0 || !tags || !mp || !nout || max_out < 0 || (max_out > 0 && !out)) return TMK_E_ARG;
    if (ncoarse < 0 || (ncoarse > 0 && !coarse_boxes)) return TMK_E_ARG;
    const int r = mp->ref_ratio, bf = mp->block_factor, mg = mp->max_grid_size;
    if (r < 2 || bf < 1 || bf % r != 0 || mg < bf || mg % bf != 0 || !(mp->fill_ratio > 0.0 && mp->fill_ratio <= 1.0))
        return TMK_E_ARG;
    const int cf = bf / r;  // tags are clustered in blocks of cf x cf cells of this level
    const int nb = mp->nest_buffer > 0 ? mp->nest_buffer : (TMK_GHOST + r - 1) / r + 1;
    // cells of this level that keep a finer box properly nested: covered by the level, with
    // every cell within nb of them inside the domain covered too
    std::vector<unsigned char> cov((size_t)nx * ny, ncoarse == 0 ? 1 : 0);
    for (int k = 0; k < ncoarse; ++k) {
        const tmk_box& b = coarse_boxes[k];
        for (int j = std::max(b.lo[1], 0); j <= std::min(b.hi[1], ny - 1); ++j)
            for (int i = std::max(b.lo[0], 0); i <= std::min(b.hi[0], nx - 1); ++i) cov[(size_t)j * nx + i] = 1;
    }
    std::vector<unsigned char> pnd((size_t)nx * ny, 1);
    if (ncoarse > 0) {
        // separable erosion: a cell stays when the (2 nb + 1)^2 neighbourhood, clipped to the
        // domain, is covered
        std::vector<unsigned char> tmp((size_t)nx * ny, 0);
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                bool ok = true;
                for (int ii = std::max(i - nb, 0); ii <= std::min(i + nb, nx - 1) && ok; ++ii)
                    ok = cov[(size_t)j * nx + ii] != 0;
                tmp[(size_t)j * nx + i] = ok;
            }
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                bool ok = true;
                for (int jj = std::max(j - nb, 0); jj <= std::min(j + nb, ny - 1) && ok; ++jj)
                    ok = tmp[(size_t)jj * nx + i] != 0;
                pnd[(size_t)j * nx + i] = ok;
            }
    }
    // tags and nesting domain coarsened by cf (a block is tagged when any of its cells is; it
    // is properly nested when all of its cells are)
    TagField g;
    g.nx = (nx + cf - 1) / cf;
    g.ny = (ny + cf - 1) / cf;
    g.t.assign((size_t)g.nx * g.ny, 0);
    std::vector<unsigned char> gp((size_t)g.nx * g.ny, 1);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const size_t q = (size_t)(j / cf) * g.nx + i / cf;
            if (tags[(size_t)j * nx + i]) g.t[q] = 1;
            if (!pnd[(size_t)j * nx + i]) gp[q] = 0;
        }
    for (size_t q = 0; q < g.t.size(); ++q)
        if (!gp[q]) g.t[q] = 0;
    std::vector<tmk_box> cl;
    br_cluster(g, tmk_box{{0, 0}, {g.nx - 1, g.ny - 1}}, mp->fill_ratio, cl);
    // boxes reaching outside the nesting domain are cut until they lie inside it
    std::vector<tmk_box> nested;
    std::function<void(tmk_box)> nest = [&](tmk_box b) {
        bool inside = true, any = false;
        for (int j = b.lo[1]; j <= b.hi[1] && inside; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i) {
                if (!gp[(size_t)j * g.nx + i]) { inside = false; break; }
            }
        for (int j = b.lo[1]; j <= b.hi[1] && !any; ++j)
            for (int i = b.lo[0]; i <= b.hi[0] && !any; ++i) any = g.at(i, j);
        if (!any) return;
        if (inside) { nested.push_back(b); return; }
        const int w = b.hi[0] - b.lo[0] + 1, h = b.hi[1] - b.lo[1] + 1;
        const int d = w >= h ? 0 : 1, k = (d == 0 ? w : h) / 2;
        tmk_box a = b, c = b;
        a.hi[d] = b.lo[d] + k - 1;
        c.lo[d] = b.lo[d] + k;
        // tight boxes around the tags of each half
        std::vector<tmk_box> sub;
        br_cluster(g, a, mp->fill_ratio, sub);
        br_cluster(g, c, mp->fill_ratio, sub);
        for (const tmk_box& x : sub) nest(x);
    };
    for (const tmk_box& b : cl) nest(b);
    // to finer-level cells, then pieces of at most max_grid_size, multiples of block_factor
    std::vector<tmk_box> res;
    const int fx = cf * r;  // finer cells per clustered block
    const int fnx = nx * r, fny = ny * r;
    for (const tmk_box& b : nested) {
        tmk_box f{{b.lo[0] * fx, b.lo[1] * fx}, {std::min((b.hi[0] + 1) * fx, fnx) - 1, std::min((b.hi[1] + 1) * fx, fny) - 1}};
        const int nbx = (f.hi[0] - f.lo[0] + 1 + mg - 1) / mg, nby = (f.hi[1] - f.lo[1] + 1 + mg - 1) / mg;
        const int blk_x = (f.hi[0] - f.lo[0] + bf) / bf, blk_y = (f.hi[1] - f.lo[1] + bf) / bf;
        for (int py = 0; py < nby; ++py)
            for (int px = 0; px < nbx; ++px) {
                tmk_box p;
                p.lo[0] = f.lo[0] + (blk_x * px / nbx) * bf;
                p.hi[0] = std::min(f.lo[0] + (blk_x * (px + 1) / nbx) * bf - 1, f.hi[0]);
                p.lo[1] = f.lo[1] + (blk_y * py / nby) * bf;
                p.hi[1] = std::min(f.lo[1] + (blk_y * (py + 1) / nby) * bf - 1, f.hi[1]);
                if (!box_empty(p)) res.push_back(p);
            }
    }
    *nout = (int)res.size();
    if ((int)res.size() > max_out) return TMK_E_OVERFLOW;
    for (size_t k = 0; k < res.size(); ++k) out[k] = res[k];
    return TMK_OK;
}

template <typename T>
static int level_regrid_t(tmk_level* N, tmk_level* O, tmk_level* C, int r)
{
    PairMaps m;
    int e = pair_maps(C, N, r, m);
    if (e) return e;
    const int ncx = m.ncx, ncy = m.ncy;
    hipStream_t st = N->ctx->stream;
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipStreamSynchronize(C->ctx->stream));
    if (O) HIPCHK(hipStreamSynchronize(O->ctx->stream));
    auto in_c = [&](int i, int j) { return i >= 0 && i < ncx && j >= 0 && j < ncy && m.c_at(i, j) >= 0; };
    const int nN = (int)N->boxes.size();
    std::vector<CfCell> cells;
    std::vector<CfFace> faces;
    std::vector<CopyRect> rects, brects;
    for (int b = 0; b < nN; ++b) {
        const tmk_box v = N->boxes[b];
        const int w = v.hi[0] - v.lo[0] + 1, h = v.hi[1] - v.lo[1] + 1;
        std::vector<unsigned char> old((size_t)w * h, 0);
        std::vector<unsigned char> oldf(N->has_b[b] ? w : 0, 0);
        if (O)
            for (int o = 0; o < (int)O->boxes.size(); ++o) {
                const tmk_box x = box_and(v, O->boxes[o]);
                if (box_empty(x)) continue;
                rects.push_back(CopyRect{N->npatch + o, b, x.lo[0], x.lo[1], x.hi[0], x.hi[1]});
                for (int j = x.lo[1]; j <= x.hi[1]; ++j)
                    for (int i = x.lo[0]; i <= x.hi[0]; ++i) old[(size_t)(i - v.lo[0]) + (size_t)(j - v.lo[1]) * w] = 1;
                if (N->has_b[b] && O->has_b[o]) {
                    brects.push_back(CopyRect{N->npatch + o, b, x.lo[0], 0, x.hi[0], 0});
                    for (int i = x.lo[0]; i <= x.hi[0]; ++i) oldf[i - v.lo[0]] = 1;
                }
            }
        // FineInterp of the cells the old layout does not cover: central slopes inside the
        // domain, one-sided at its edges, limited off the domain's edge layer
        for (int j = v.lo[1]; j <= v.hi[1]; ++j)
            for (int i = v.lo[0]; i <= v.hi[0]; ++i) {
                if (old[(size_t)(i - v.lo[0]) + (size_t)(j - v.lo[1]) * w]) continue;
                const int ic = i / r, jc = j / r;
                CfCell cc;
                memset(&cc, 0, sizeof cc);
                cc.dpatch = b;
                cc.doff = cell_off(N, b, i, j);
                for (int jj = -1, k = 0; jj <= 1; ++jj)
                    for (int ii = -1; ii <= 1; ++ii, ++k) {
                        const int ci = ic + ii, cj = jc + jj;
                        cc.cpatch[k] = -1;
                        if (in_c(ci, cj)) {
                            const int cb = m.c_at(ci, cj);
                            cc.cpatch[k] = cb;
                            cc.coff[k] = cell_off(C, cb, ci, cj);
                        }
                    }
                const int kx = ic == 0 ? 2 : ic == ncx - 1 ? 1 : 0, ky = jc == 0 ? 2 : jc == ncy - 1 ? 1 : 0;
                const bool lim = ic >= 1 && ic <= ncx - 2 && jc >= 1 && jc <= ncy - 2;
                if (cc.cpatch[4] < 0 || (kx != 1 && cc.cpatch[5] < 0) || (kx != 2 && cc.cpatch[3] < 0) ||
                    (ky != 1 && cc.cpatch[7] < 0) || (ky != 2 && cc.cpatch[1] < 0))
                    return TMK_E_NESTING;
                if (lim)
                    for (int k = 0; k < 9; ++k)
                        if (cc.cpatch[k] < 0) return TMK_E_NESTING;
                cc.flags = kx | (ky << 2) | (lim ? 16 : 0);
                cc.coef0 = -0.5 + ((double)(i - r * ic) + 0.5) / (double)r;
                cc.coef1 = -0.5 + ((double)(j - r * jc) + 0.5) / (double)r;
                cells.push_back(cc);
            }
        if (!N->has_b[b]) continue;
        for (int i = v.lo[0]; i <= v.hi[0]; ++i) {
            if (oldf[i - v.lo[0]]) continue;
            const int ic = i / r;
            CfFace ff;
            memset(&ff, 0, sizeof ff);
            ff.dpatch = b;
            ff.doff = face_off(N, b, i);
            for (int k = 0; k < 3; ++k) {
                const int ci = ic - 1 + k;
                const int cb = ci >= 0 && ci < ncx ? m.own_cb[ci] : -1;
                ff.cpatch[k] = cb;
                ff.coff[k] = cb >= 0 ? face_off(C, cb, ci) : 0;
            }
            ff.kind = ic == 0 ? 2 : ic == ncx - 1 ? 1 : 0;
            ff.limit = ic >= 1 && ic <= ncx - 2;
            if (ff.cpatch[1] < 0 || (ff.kind != 1 && ff.cpatch[2] < 0) || (ff.kind != 2 && ff.cpatch[0] < 0))
                return TMK_E_NESTING;
            ff.coef = -0.5 + ((double)(i - r * ic) + 0.5) / (double)r;
            faces.push_back(ff);
        }
    }
    std::vector<int> comps;
    for (int k = 0; k < TMK_NCOMP; ++k) comps.push_back(k);
    // descriptor tables: the new level and the coarse one with .in = their current data; the
    // copy table (new level, then the old one) with .out = current data, as k_exchange reads it
    const size_t ps = sizeof(Patch<T>);
    std::vector<char> copytab(N->hpatch[N->cur ^ 1]);
    if (O) copytab.insert(copytab.end(), O->hpatch[O->cur ^ 1].begin(), O->hpatch[O->cur ^ 1].end());
    DevBuf dN, dC, dX, dcells, dfaces, dcomps, drects, dbrects;
    if ((e = upload_list(dN, N->hpatch[N->cur], st)) || (e = upload_list(dC, C->hpatch[C->cur], st)) ||
        (e = upload_list(dX, copytab, st)) || (e = upload_list(dcells, cells, st)) ||
        (e = upload_list(dfaces, faces, st)) || (e = upload_list(dcomps, comps, st)) ||
        (e = upload_list(drects, rects, st)) || (e = upload_list(dbrects, brects, st)))
        return e;
    (void)ps;
    const Patch<T>* tN = (const Patch<T>*)dN.p;
    const Patch<T>* tC = (const Patch<T>*)dC.p;
    const int nb = N->ctx->hp.fault_kind == TMK_FAULT_RS ? TMK_NBDRY_RS : TMK_NBDRY_SW;
    if (!cells.empty()) {
        hipLaunchKernelGGL(k_amr_fill_state<T>, amr_grid((long long)cells.size() * TMK_NCOMP), dim3(256), 0, st,
                           (const CfCell*)dcells.p, (int)cells.size(), (const int*)dcomps.p, TMK_NCOMP, tN, tC, tC,
                           (const Scen*)nullptr, (const Scen*)nullptr, r, 1.0);
        HIPCHK(hipGetLastError());
    }
    if (!faces.empty()) {
        hipLaunchKernelGGL(k_amr_fill_fault<T>, amr_grid((long long)faces.size() * nb), dim3(256), 0, st,
                           (const CfFace*)dfaces.p, (int)faces.size(), nb, tN, tC, tC, (const Scen*)nullptr,
                           (const Scen*)nullptr, r, 1.0);
        HIPCHK(hipGetLastError());
    }
    if (!rects.empty()) {
        hipLaunchKernelGGL(k_exchange<T>, dim3((unsigned)rects.size()), dim3(128), 0, st, (const Patch<T>*)dX.p,
                           (const CopyRect*)drects.p, (int)rects.size(), (1 << TMK_NCOMP) - 1);
        HIPCHK(hipGetLastError());
    }
    if (!brects.empty()) {
        hipLaunchKernelGGL(k_bexchange<T>, dim3((unsigned)brects.size()), dim3(64), 0, st, (const Patch<T>*)dX.p,
                           (const CopyRect*)dbrects.p, (int)brects.size());
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(st));
    N->dirty = true;
    return TMK_OK;
}

extern "C" int tmk_level_regrid(tmk_level* fresh, tmk_level* old, tmk_level* coarse, int ref_ratio)
{
    if (!fresh || !coarse || ref_ratio < 2 || ref_ratio > TMK_AMR_MAXR || fresh == coarse || old == fresh) return TMK_E_ARG;
    for (tmk_level* L : {fresh, old, coarse})
        if (L && (L->batch || L->amr_stream)) return TMK_E_ARG;
    if (fresh->tsize != coarse->tsize || (old && old->tsize != fresh->tsize) || fresh->ns != coarse->ns ||
        (old && old->ns != fresh->ns))
        return TMK_E_ARG;
    if (old && (old->ctx->hp.domain.hi[0] != fresh->ctx->hp.domain.hi[0] ||
                old->ctx->hp.domain.hi[1] != fresh->ctx->hp.domain.hi[1]))
        return TMK_E_ARG;
    return fresh->tsize == sizeof(double) ? level_regrid_t<double>(fresh, old, coarse, ref_ratio)
                                          : level_regrid_t<float>(fresh, old, coarse, ref_ratio);
}
