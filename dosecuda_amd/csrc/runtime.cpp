// runtime.cpp -- per-device state (stream, scratch arena), the tuning table
// and launch timing.
#include "dh_runtime.h"

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

struct ws_slots {
    void *ptr[DH_SLOT_COUNT_] = {};
    size_t bytes[DH_SLOT_COUNT_] = {};
};

struct dev_state {
    bool ready = false;
    hipStream_t stream = nullptr;
    ws_slots ws[DH_MAX_WORKSPACES];
};

constexpr int kMaxDevices = 64;
dev_state g_dev[kMaxDevices];
std::recursive_mutex g_mu;
thread_local int t_ws = 0;

struct prof_rec {
    std::string name;
    hipEvent_t a, b;
};
bool g_prof_on = false;
std::vector<prof_rec> g_prof;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t take_event()
{
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    DH_HIP(hipEventCreate(&e));
    return e;
}

}  // namespace

int dh_current_device()
{
    int d = 0;
    DH_HIP(hipGetDevice(&d));
    return d;
}

void dh_device_enter(int dev)
{
    int n = 0;
    DH_HIP(hipGetDeviceCount(&n));
    if (dev < 0 || dev >= n || dev >= kMaxDevices) {
        throw dh_error(DH_E_RUNTIME, hipGetErrorString(hipErrorInvalidDevice));
    }
    DH_HIP(hipSetDevice(dev));
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    dev_state &st = g_dev[dev];
    if (!st.ready) {
        DH_HIP(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
        st.ready = true;
    }
}

hipStream_t dh_stream(int dev)
{
    dh_device_enter(dev);
    return g_dev[dev].stream;
}

dh_side &dh_side_of(int dev)
{
    static std::mutex mu;
    static dh_side side[kMaxDevices];
    DH_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    dh_side &sd = side[dev];
    if (!sd.st) {
        DH_HIP(hipStreamCreateWithFlags(&sd.st, hipStreamNonBlocking));
        DH_HIP(hipEventCreateWithFlags(&sd.fork, hipEventDisableTiming));
        DH_HIP(hipEventCreateWithFlags(&sd.join, hipEventDisableTiming));
    }
    return sd;
}

void dh_set_workspace(int ws)
{
    if (ws < 0 || ws >= DH_MAX_WORKSPACES) throw dh_error(DH_E_INVALID, "workspace index out of range");
    t_ws = ws;
}

int dh_get_workspace() { return t_ws; }

void *dh_scratch(int dev, int slot, size_t bytes)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    ws_slots &st = g_dev[dev].ws[t_ws];
    if (bytes == 0) bytes = 256;
    if (st.bytes[slot] < bytes) {
        if (st.ptr[slot]) {
            DH_HIP(hipDeviceSynchronize());
            DH_HIP(hipFree(st.ptr[slot]));
            st.ptr[slot] = nullptr;
            st.bytes[slot] = 0;
        }
        size_t b = (bytes + 4095) & ~(size_t)4095;
        void *p = nullptr;
        DH_HIP(hipMalloc(&p, b));
        st.ptr[slot] = p;
        st.bytes[slot] = b;
    }
    return st.ptr[slot];
}

void dh_device_release(int dev)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    dev_state &st = g_dev[dev];
    if (!st.ready) return;
    DH_HIP(hipSetDevice(dev));
    DH_HIP(hipDeviceSynchronize());
    for (int w = 0; w < DH_MAX_WORKSPACES; w++)
        for (int s = 0; s < DH_SLOT_COUNT_; s++) {
            if (st.ws[w].ptr[s]) DH_HIP(hipFree(st.ws[w].ptr[s]));
            st.ws[w].ptr[s] = nullptr;
            st.ws[w].bytes[s] = 0;
        }
}

// ---------------------------------------------------------------------------
// tuning table
// ---------------------------------------------------------------------------
namespace {
struct tune_entry {
    std::string kernel, key;
    int choice;
    double ms;
};
std::vector<tune_entry> g_tune;
std::mutex g_tune_mu;
}  // namespace

int dh_tune_lookup(const char *kernel, const std::string &key, int dflt)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (const auto &e : g_tune)
        if (e.kernel == kernel && e.key == key) return e.choice;
    return dflt;
}

void dh_tune_store(const char *kernel, const std::string &key, int choice, double ms)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (auto &e : g_tune)
        if (e.kernel == kernel && e.key == key) {
            e.choice = choice;
            e.ms = ms;
            return;
        }
    g_tune.push_back(tune_entry{kernel, key, choice, ms});
}

std::string dh_tune_key_geom(const dh_geom &g)
{
    const float a0 = std::fabs(g.src[0]), a1 = std::fabs(g.src[1]), a2 = std::fabs(g.src[2]);
    const int axis = (a0 >= a1 && a0 >= a2) ? 0 : (a1 >= a2 ? 1 : 2);
    return std::to_string(g.nz) + "," + std::to_string(g.ny) + "," + std::to_string(g.nx) + "," +
           std::to_string(axis);
}

dh_tile dh_make_tile(int choice, int nx, int ny, int nz, unsigned *grid)
{
    // choices 8-11: bricks of choices 3, 5, 6 and 4 numbered y fastest
    static const int shapes[DH_TILE_CHOICES][4] = {{0, 0, 0, 0},  {64, 4, 1, 0}, {32, 8, 1, 0}, {16, 16, 1, 0},
                                                   {32, 4, 2, 0}, {16, 4, 4, 0}, {8, 8, 4, 0},  {64, 2, 2, 0},
                                                   {16, 16, 1, 1}, {16, 4, 4, 1}, {8, 8, 4, 1}, {32, 4, 2, 1}};
    DH_REQUIRE(choice >= 0 && choice < DH_TILE_CHOICES, "tile choice out of range");
    dh_tile t;
    t.tx = shapes[choice][0];
    t.ty = shapes[choice][1];
    t.tz = shapes[choice][2];
    t.yfirst = shapes[choice][3];
    if (t.tx == 0) {
        t.ntx = t.nty = 0;
        *grid = (unsigned)(((size_t)nx * ny * nz + 255) / 256);
    } else {
        t.ntx = (nx + t.tx - 1) / t.tx;
        t.nty = (ny + t.ty - 1) / t.ty;
        *grid = (unsigned)((size_t)t.ntx * t.nty * ((nz + t.tz - 1) / t.tz));
    }
    return t;
}

extern "C" int dh_tune_set(const char *kernel, const char *key, int32_t choice, double ms)
{
    return dh_guarded([&] {
        DH_REQUIRE(kernel && key, "null argument");
        dh_tune_store(kernel, key, choice, ms);
    });
}

extern "C" int dh_tune_get(const char *kernel, const char *key, int32_t *choice)
{
    return dh_guarded([&] {
        DH_REQUIRE(kernel && key && choice, "null argument");
        *choice = dh_tune_lookup(kernel, key, -1);
    });
}

extern "C" int dh_tune_count(int32_t *n)
{
    return dh_guarded([&] {
        DH_REQUIRE(n, "null argument");
        std::lock_guard<std::mutex> lk(g_tune_mu);
        *n = (int32_t)g_tune.size();
    });
}

extern "C" int dh_tune_entry(int32_t index, char *kernel, int32_t kernel_cap, char *key, int32_t key_cap,
                             int32_t *choice, double *ms)
{
    return dh_guarded([&] {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        DH_REQUIRE(index >= 0 && index < (int32_t)g_tune.size(), "entry index out of range");
        const tune_entry &e = g_tune[index];
        DH_REQUIRE(kernel && key && kernel_cap > (int32_t)e.kernel.size() && key_cap > (int32_t)e.key.size(),
                   "name buffers too small");
        std::memcpy(kernel, e.kernel.c_str(), e.kernel.size() + 1);
        std::memcpy(key, e.key.c_str(), e.key.size() + 1);
        if (choice) *choice = e.choice;
        if (ms) *ms = e.ms;
    });
}

extern "C" int dh_tune_clear(void)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune.clear();
    return DH_OK;
}

// ---------------------------------------------------------------------------
// launch timing
// ---------------------------------------------------------------------------
dh_timed::dh_timed(const char *n, hipStream_t st) : name(n), s(st), ev0(nullptr)
{
    if (!g_prof_on) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    hipEvent_t e = take_event();
    DH_HIP(hipEventRecord(e, s));
    ev0 = (void *)e;
}

dh_timed::~dh_timed()
{
    if (!ev0) return;
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    hipEvent_t e = take_event();
    if (hipEventRecord(e, s) != hipSuccess) return;
    g_prof.push_back(prof_rec{name, (hipEvent_t)ev0, e});
}

extern "C" int dh_prof_enable(int on)
{
    g_prof_on = on != 0;
    return DH_OK;
}

extern "C" int dh_prof_reset(void)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!g_prof.empty()) (void)hipDeviceSynchronize();
    for (auto &r : g_prof) {
        g_event_pool.push_back(r.a);
        g_event_pool.push_back(r.b);
    }
    g_prof.clear();
    return DH_OK;
}

extern int dh_set_error(int code, const char *msg);

extern "C" int dh_prof_get(const char *kernel, double *total_ms, int64_t *launches)
{
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    double tot = 0.0;
    int64_t n = 0;
    for (auto &r : g_prof) {
        if (kernel && std::strcmp(kernel, r.name.c_str()) != 0) continue;
        if (hipEventSynchronize(r.b) != hipSuccess) return dh_set_error(DH_E_RUNTIME, "event sync failed");
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) return dh_set_error(DH_E_RUNTIME, "event timing failed");
        tot += ms;
        n++;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = n;
    return DH_OK;
}
