// scheduler.hip -- batching scheduler for many independent dose cases on one
// GPU.  A case is one phantom (grid + input volume on the device) with its
// jobs: photon control points (plain CCCS or the checked surrogate), proton
// WET, proton dense dose, influence-matrix fills.  The jobs of a case run in
// order on one stream and accumulate into the case's dose; cases are spread
// over the scheduler's streams, longest first onto the least-loaded stream
// (LPT), by measured per-job GPU times once an eager run has timed them, by a
// voxel-count estimate before.  Each stream has its own scratch workspace, so
// concurrent cases never share buffers.  A batch can be captured once into a
// single HIP graph (fork/join over the streams) and replayed.
//
// No reference counterpart: DoseCUDA runs one synchronous call per beam
// (dosemodule.cu:95-538).
#include "dh_runtime.h"

#include <cstring>

#include <cmath>

#include <algorithm>
#include <memory>
#include <numeric>
#include <vector>

namespace {

__global__ void k_sched_fill(float *__restrict__ p, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

// dst += src * scale: the product rounded to float32, then the sum (as a
// float32 array times a float32 factor added to another); scale 1 is exact.
__global__ void k_sched_accum(float *__restrict__ dst, const float *__restrict__ src, int64_t n, float scale)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float p = __fmul_rn(src[i], scale);
        dst[i] = __fadd_rn(dst[i], p);
    }
}

unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); }

enum job_kind { J_PHOTON = 0, J_PROTON_WET, J_PROTON_DOSE, J_INFL_FILL };

// deep copy of a photon beam model (the C struct points at caller arrays)
struct model_copy {
    dh_photon_model m;
    std::vector<float> pr, pi, ps, mu, pw, sw;
    void set(const dh_photon_model &src)
    {
        DH_REQUIRE(src.n_profile > 0 && src.profile_radius && src.profile_intensities && src.profile_softening,
                   "bad photon model profile");
        DH_REQUIRE(src.n_spectral >= 0, "bad photon model spectrum");
        m = src;
        pr.assign(src.profile_radius, src.profile_radius + src.n_profile);
        pi.assign(src.profile_intensities, src.profile_intensities + src.n_profile);
        ps.assign(src.profile_softening, src.profile_softening + src.n_profile);
        if (src.n_spectral) {
            DH_REQUIRE(src.spectrum_attenuation_coefficients && src.spectrum_primary_weights &&
                       src.spectrum_scatter_weights, "bad photon model spectrum");
            mu.assign(src.spectrum_attenuation_coefficients, src.spectrum_attenuation_coefficients + src.n_spectral);
            pw.assign(src.spectrum_primary_weights, src.spectrum_primary_weights + src.n_spectral);
            sw.assign(src.spectrum_scatter_weights, src.spectrum_scatter_weights + src.n_spectral);
        }
        m.profile_radius = pr.data();
        m.profile_intensities = pi.data();
        m.profile_softening = ps.data();
        m.spectrum_attenuation_coefficients = mu.empty() ? nullptr : mu.data();
        m.spectrum_primary_weights = pw.empty() ? nullptr : pw.data();
        m.spectrum_scatter_weights = sw.empty() ? nullptr : sw.data();
    }
};

struct job {
    int kind = J_PHOTON;
    int case_id = 0;
    // photon
    dh_photon_cp cp{};
    model_copy model;
    std::vector<float> mlc;
    const dh_sur_net *net = nullptr;
    int stride = 3;
    float tol = 0.01f;
    int dilate = 2;
    float scale = 1.0f;          // factor of the photon dose before it is added
    // proton
    dh_proton_beam beam{};
    float *wet_out = nullptr;
    dh_proton_plan *plan = nullptr;
    const float *wet_in = nullptr;
    // influence
    dh_influence *infl = nullptr;
    uint32_t *col = nullptr;
    float *val = nullptr;
    // timing
    double est = 0.0, ms = -1.0;
    hipEvent_t ev[2] = {nullptr, nullptr};
};

struct scase {
    dh_grid grid{};
    const float *vol = nullptr;
    float *dose = nullptr;
    std::vector<int> jobs;
    int stream = 0;
};

}  // namespace

struct dh_sched {
    int gpu = 0;
    std::vector<hipStream_t> streams;
    hipStream_t origin = nullptr;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr, done = nullptr;
    std::vector<scase> cases;
    std::vector<std::unique_ptr<job>> jobs;
    bool timed = false;          // per-job times measured by an eager run
    bool planned = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double last_ms = -1.0;
    hipEvent_t t0 = nullptr, t1 = nullptr;
};

namespace {

int64_t nvox(const dh_grid &g) { return (int64_t)g.nx * g.ny * g.nz; }

void drop_graph(dh_sched *s)
{
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->exec = nullptr;
    s->graph = nullptr;
}

// LPT: cases by decreasing cost onto the least-loaded stream
void plan_streams(dh_sched *s)
{
    const int ns = (int)s->streams.size();
    std::vector<double> cost(s->cases.size(), 0.0);
    for (size_t c = 0; c < s->cases.size(); c++)
        for (int j : s->cases[c].jobs) {
            const job &J = *s->jobs[j];
            cost[c] += (s->timed && J.ms >= 0.0) ? J.ms : J.est;
        }
    std::vector<int> order(s->cases.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::vector<double> load(ns, 0.0);
    for (int c : order) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        s->cases[c].stream = k;
        load[k] += cost[c];
    }
    s->planned = true;
}

// Two photon control points with the same isocentre, gantry and couch angle
// and source distance have the same radiological depth (it does not depend
// on the collimator, the MLC or the MU).
bool same_source(const job &a, const job &b)
{
    return a.kind == J_PHOTON && b.kind == J_PHOTON && a.cp.iso[0] == b.cp.iso[0] && a.cp.iso[1] == b.cp.iso[1] &&
           a.cp.iso[2] == b.cp.iso[2] && a.cp.ga == b.cp.ga && a.cp.ta == b.cp.ta &&
           a.model.m.primary_source_distance == b.model.m.primary_source_distance;
}

void enqueue_job(dh_sched *s, job &J, scase &C, hipStream_t st, bool timing, bool reuse_wet)
{
    if (timing) DH_HIP(hipEventRecord(J.ev[0], st));
    const int64_t n = nvox(C.grid);
    switch (J.kind) {
    case J_PHOTON: {
        float *tmp = (float *)dh_scratch(s->gpu, DH_SLOT_OUT1, sizeof(float) * (size_t)n);
        dh_photon_enqueue(s->gpu, &C.grid, &J.cp, &J.model.m, J.mlc.data(), (int)(J.mlc.size() / 4), C.vol, tmp,
                          J.net, J.stride, J.tol, J.dilate, st, reuse_wet);
        hipLaunchKernelGGL(k_sched_accum, dim3(grid_of(n)), dim3(256), 0, st, C.dose, (const float *)tmp, n,
                           J.scale);
        break;
    }
    case J_PROTON_WET:
        dh_proton_wet_enqueue(s->gpu, &C.grid, &J.beam, C.vol, J.wet_out, st);
        break;
    case J_PROTON_DOSE: {
        float *tmp = (float *)dh_scratch(s->gpu, DH_SLOT_OUT1, sizeof(float) * (size_t)n);
        dh_proton_plan_enqueue(J.plan, J.wet_in, tmp, st);
        hipLaunchKernelGGL(k_sched_accum, dim3(grid_of(n)), dim3(256), 0, st, C.dose, (const float *)tmp, n, 1.0f);
        break;
    }
    case J_INFL_FILL:
        dh_influence_fill_enqueue(J.infl, J.wet_in, J.col, J.val, st);
        break;
    }
    DH_HIP(hipGetLastError());
    if (timing) DH_HIP(hipEventRecord(J.ev[1], st));
}

// The same beam model: every scalar field and table equal.
bool same_model(const model_copy &a, const model_copy &b)
{
    dh_photon_model x = a.m, y = b.m;
    for (dh_photon_model *m : {&x, &y}) {
        m->profile_radius = m->profile_intensities = m->profile_softening = nullptr;
        m->spectrum_attenuation_coefficients = m->spectrum_primary_weights = m->spectrum_scatter_weights = nullptr;
    }
    return std::memcmp(&x, &y, sizeof(x)) == 0 && a.pr == b.pr && a.pi == b.pi && a.ps == b.ps && a.mu == b.mu &&
           a.pw == b.pw && a.sw == b.sw;
}

// Two checked-surrogate photon jobs of one case that can share a network
// launch: the same network, check settings and beam model.
bool same_sur_group(const job &a, const job &b)
{
    if (a.kind != J_PHOTON || b.kind != J_PHOTON || !a.net || a.net != b.net) return false;
    if (a.stride != b.stride || a.tol != b.tol || a.dilate != b.dilate) return false;
    return same_model(a.model, b.model);
}

// Up to kSurGroupMax consecutive surrogate photon jobs J[0..n) of case C as
// one batched network and check (dh_photon_sur_group_enqueue), each dose
// then added with its own scale.  The group's time is the first job's; the
// others record empty intervals after it.
void enqueue_sur_group(dh_sched *s, job *const *J, int n, scase &C, hipStream_t st, bool timing, bool reuse_first)
{
    if (timing) DH_HIP(hipEventRecord(J[0]->ev[0], st));
    const int64_t nv = nvox(C.grid);
    float *tmp = (float *)dh_scratch(s->gpu, DH_SLOT_OUT1, sizeof(float) * (size_t)(nv * n));
    std::vector<dh_photon_cp> cps(n);
    std::vector<const float *> mlc(n);
    std::vector<int> np(n);
    std::vector<float *> out(n);
    for (int i = 0; i < n; i++) {
        cps[i] = J[i]->cp;
        mlc[i] = J[i]->mlc.data();
        np[i] = (int)(J[i]->mlc.size() / 4);
        out[i] = tmp + (size_t)nv * i;
    }
    const job &J0 = *J[0];
    dh_photon_sur_group_enqueue(s->gpu, &C.grid, n, cps.data(), &J0.model.m, mlc.data(), np.data(), C.vol, out.data(),
                                J0.net, J0.stride, J0.tol, J0.dilate, st, reuse_first);
    for (int i = 0; i < n; i++)
        hipLaunchKernelGGL(k_sched_accum, dim3(grid_of(nv)), dim3(256), 0, st, C.dose, (const float *)out[i], nv,
                           J[i]->scale);
    DH_HIP(hipGetLastError());
    if (timing) {
        DH_HIP(hipEventRecord(J[0]->ev[1], st));
        for (int i = 1; i < n; i++) {
            DH_HIP(hipEventRecord(J[i]->ev[0], st));
            DH_HIP(hipEventRecord(J[i]->ev[1], st));
        }
    }
}

// all cases' work on the worker streams, after `fork` and before `join`
void enqueue_batch(dh_sched *s, bool timing)
{
    const int ws0 = dh_get_workspace();
    try {
        for (size_t k = 0; k < s->streams.size(); k++) DH_HIP(hipStreamWaitEvent(s->streams[k], s->fork, 0));
        for (auto &C : s->cases) {
            hipStream_t st = s->streams[C.stream];
            dh_set_workspace(1 + C.stream);
            if (C.dose) {
                const int64_t n = nvox(C.grid);
                hipLaunchKernelGGL(k_sched_fill, dim3(grid_of(n)), dim3(256), 0, st, C.dose, n);
            }
            // a case's jobs run back to back on its stream and workspace, so a
            // control point may reuse the ray trace of the one before it;
            // consecutive checked-surrogate control points share network
            // launches in groups of up to kSurGroupMax
            const job *prev = nullptr;
            for (size_t q = 0; q < C.jobs.size();) {
                job &J = *s->jobs[C.jobs[q]];
                const bool reuse = prev && same_source(*prev, J);
                if (J.kind == J_PHOTON && J.net) {
                    job *G[kSurGroupMax];
                    int n = 0;
                    while (q + n < C.jobs.size() && n < kSurGroupMax &&
                           same_sur_group(J, *s->jobs[C.jobs[q + n]])) {
                        G[n] = s->jobs[C.jobs[q + n]].get();
                        n++;
                    }
                    enqueue_sur_group(s, G, n, C, st, timing, reuse);
                    prev = G[n - 1];
                    q += n;
                    continue;
                }
                enqueue_job(s, J, C, st, timing, reuse);
                prev = &J;
                q++;
            }
        }
        for (size_t k = 0; k < s->streams.size(); k++) {
            DH_HIP(hipEventRecord(s->join[k], s->streams[k]));
            DH_HIP(hipStreamWaitEvent(s->origin, s->join[k], 0));
        }
    } catch (...) {
        dh_set_workspace(ws0);
        throw;
    }
    dh_set_workspace(ws0);
}

}  // namespace

extern "C" int dh_sched_create(int gpu_id, int32_t n_streams, dh_sched **out)
{
    return dh_guarded([&] {
        DH_REQUIRE(out, "null output handle");
        DH_REQUIRE(n_streams >= 1 && n_streams <= DH_SCHED_WORKSPACES, "1..32 streams");
        dh_device_enter(gpu_id);
        std::unique_ptr<dh_sched> s(new dh_sched());
        s->gpu = gpu_id;
        for (int k = 0; k < n_streams; k++) {
            hipStream_t st;
            DH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            s->streams.push_back(st);
            hipEvent_t e;
            DH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            s->join.push_back(e);
        }
        DH_HIP(hipStreamCreateWithFlags(&s->origin, hipStreamNonBlocking));
        DH_HIP(hipEventCreateWithFlags(&s->fork, hipEventDisableTiming));
        DH_HIP(hipEventCreateWithFlags(&s->done, hipEventDisableTiming));
        DH_HIP(hipEventCreate(&s->t0));
        DH_HIP(hipEventCreate(&s->t1));
        *out = s.release();
    });
}

extern "C" int dh_sched_destroy(dh_sched *s)
{
    if (!s) return DH_OK;
    (void)hipSetDevice(s->gpu);
    (void)hipStreamSynchronize(s->origin);
    for (auto st : s->streams) (void)hipStreamSynchronize(st);
    drop_graph(s);
    for (auto &J : s->jobs)
        for (auto e : J->ev)
            if (e) (void)hipEventDestroy(e);
    for (auto e : s->join) (void)hipEventDestroy(e);
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    (void)hipEventDestroy(s->fork);
    (void)hipEventDestroy(s->done);
    (void)hipEventDestroy(s->t0);
    (void)hipEventDestroy(s->t1);
    (void)hipStreamDestroy(s->origin);
    delete s;
    return DH_OK;
}

extern "C" int dh_sched_add_case(dh_sched *s, const dh_grid *grid, const float *d_volume, float *d_dose,
                                 int32_t *case_id)
{
    return dh_guarded([&] {
        DH_REQUIRE(s && grid && d_volume && case_id, "null argument");
        DH_REQUIRE(grid->nx > 0 && grid->ny > 0 && grid->nz > 0 && grid->spacing > 0.0f, "bad grid");
        drop_graph(s);
        scase c;
        c.grid = *grid;
        c.vol = d_volume;
        c.dose = d_dose;
        s->cases.push_back(c);
        s->planned = false;
        *case_id = (int32_t)s->cases.size() - 1;
    });
}

namespace {

job &new_job(dh_sched *s, int32_t case_id, int kind, double est)
{
    DH_REQUIRE(s, "null scheduler");
    DH_REQUIRE(case_id >= 0 && case_id < (int32_t)s->cases.size(), "unknown case");
    drop_graph(s);
    std::unique_ptr<job> J(new job());
    J->kind = kind;
    J->case_id = case_id;
    J->est = est;
    DH_HIP(hipSetDevice(s->gpu));
    DH_HIP(hipEventCreate(&J->ev[0]));
    DH_HIP(hipEventCreate(&J->ev[1]));
    s->jobs.push_back(std::move(J));
    s->cases[case_id].jobs.push_back((int)s->jobs.size() - 1);
    s->planned = false;
    s->timed = false;
    return *s->jobs.back();
}

}  // namespace

extern "C" int dh_sched_add_photon_scaled(dh_sched *s, int32_t case_id, const dh_photon_cp *cp,
                                          const dh_photon_model *model, const float *mlc_rows, int32_t n_pairs,
                                          const dh_sur_net *net, int32_t stride, float tol, int32_t dilate,
                                          float scale)
{
    return dh_guarded([&] {
        DH_REQUIRE(std::isfinite(scale), "the dose scale must be finite");
        DH_REQUIRE(cp && model, "null argument");
        dh_check_dilate_word(dilate);
        DH_REQUIRE(n_pairs >= 0 && n_pairs <= DH_MAX_PAIRS && (n_pairs == 0 || mlc_rows), "bad MLC");
        DH_REQUIRE(!net || (stride >= 1 && stride <= 64), "lattice stride must be 1..64");
        DH_REQUIRE(s && case_id >= 0 && case_id < (int32_t)s->cases.size(), "unknown case");
        DH_REQUIRE(s->cases[case_id].dose, "the case has no dose output");
        const double est = 1e-6 * (double)nvox(s->cases[case_id].grid) * (net ? 0.5 : 1.0);
        job &J = new_job(s, case_id, J_PHOTON, est);
        J.cp = *cp;
        J.model.set(*model);
        J.mlc.assign(mlc_rows, mlc_rows + 4 * (size_t)n_pairs);
        J.net = net;
        J.stride = stride;
        J.tol = tol;
        J.dilate = dilate;
        J.scale = scale;
    });
}

extern "C" int dh_sched_add_photon(dh_sched *s, int32_t case_id, const dh_photon_cp *cp,
                                   const dh_photon_model *model, const float *mlc_rows, int32_t n_pairs,
                                   const dh_sur_net *net, int32_t stride, float tol, int32_t dilate)
{
    return dh_sched_add_photon_scaled(s, case_id, cp, model, mlc_rows, n_pairs, net, stride, tol, dilate, 1.0f);
}

extern "C" int dh_sched_add_proton_wet(dh_sched *s, int32_t case_id, const dh_proton_beam *beam, float *d_wet_out)
{
    return dh_guarded([&] {
        DH_REQUIRE(beam && d_wet_out, "null argument");
        DH_REQUIRE(s && case_id >= 0 && case_id < (int32_t)s->cases.size(), "unknown case");
        job &J = new_job(s, case_id, J_PROTON_WET, 0.3e-6 * (double)nvox(s->cases[case_id].grid));
        J.beam = *beam;
        J.wet_out = d_wet_out;
    });
}

extern "C" int dh_sched_add_proton_dose(dh_sched *s, int32_t case_id, dh_proton_plan *plan, const float *d_wet)
{
    return dh_guarded([&] {
        DH_REQUIRE(plan && d_wet, "null argument");
        DH_REQUIRE(s && case_id >= 0 && case_id < (int32_t)s->cases.size(), "unknown case");
        DH_REQUIRE(dh_proton_plan_device(plan) == s->gpu, "plan is on another device");
        DH_REQUIRE(s->cases[case_id].dose, "the case has no dose output");
        job &J = new_job(s, case_id, J_PROTON_DOSE, 1e-6 * (double)nvox(s->cases[case_id].grid));
        J.plan = plan;
        J.wet_in = d_wet;
    });
}

extern "C" int dh_sched_add_influence_fill(dh_sched *s, int32_t case_id, dh_influence *f, const float *d_wet,
                                           uint32_t *d_col, float *d_val)
{
    return dh_guarded([&] {
        DH_REQUIRE(f && d_wet, "null argument");
        DH_REQUIRE(s && case_id >= 0 && case_id < (int32_t)s->cases.size(), "unknown case");
        DH_REQUIRE(dh_influence_device(f) == s->gpu, "influence matrix is on another device");
        const int64_t nnz = dh_influence_nnz(f);
        DH_REQUIRE(nnz >= 0, "dh_influence_count must run before the fill is scheduled");
        DH_REQUIRE(nnz == 0 || (d_col && d_val), "null CSR arrays");
        dh_device_enter(s->gpu);
        dh_influence_fill_prepare(f, s->origin);
        job &J = new_job(s, case_id, J_INFL_FILL, 2e-8 * (double)nnz);
        J.infl = f;
        J.wet_in = d_wet;
        J.col = d_col;
        J.val = d_val;
    });
}

// mode 0: eager, every job timed with events (the next plan uses the
// measured times); mode 1: one HIP graph of the whole batch, captured on
// first use after an eager warm-up run (which sizes the workspaces) and
// replayed afterwards.  The batch starts after the work already queued on
// `stream` and that stream's later work waits for it; the call returns
// after enqueueing.
extern "C" int dh_sched_run(dh_sched *s, int32_t mode, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(s, "null scheduler");
        DH_REQUIRE(mode == 0 || mode == 1, "mode must be 0 (eager) or 1 (graph)");
        dh_device_enter(s->gpu);
        hipStream_t caller = (hipStream_t)stream;
        DH_HIP(hipEventRecord(s->done, caller));
        DH_HIP(hipStreamWaitEvent(s->origin, s->done, 0));
        if (mode == 1 && !s->exec) {
            if (!s->timed) {
                // timed warm-up on the estimated plan
                if (!s->planned) plan_streams(s);
                DH_HIP(hipEventRecord(s->fork, s->origin));
                enqueue_batch(s, true);
                DH_HIP(hipStreamSynchronize(s->origin));
                for (auto &J : s->jobs) {
                    float ms = 0.f;
                    DH_HIP(hipEventElapsedTime(&ms, J->ev[0], J->ev[1]));
                    J->ms = ms;
                }
                s->timed = true;
                s->planned = false;
            }
            // the graph uses the plan on measured times; one eager run on that
            // plan grows every workspace to its cases' sizes, so nothing is
            // allocated (or synchronised) under capture
            plan_streams(s);
            DH_HIP(hipEventRecord(s->fork, s->origin));
            enqueue_batch(s, false);
            DH_HIP(hipStreamSynchronize(s->origin));
        }
        if (!s->planned) plan_streams(s);
        if (mode == 0) {
            DH_HIP(hipEventRecord(s->t0, s->origin));
            DH_HIP(hipEventRecord(s->fork, s->origin));
            enqueue_batch(s, true);
            DH_HIP(hipEventRecord(s->t1, s->origin));
            DH_HIP(hipStreamSynchronize(s->origin));
            float ms = 0.f;
            DH_HIP(hipEventElapsedTime(&ms, s->t0, s->t1));
            s->last_ms = ms;
            for (auto &J : s->jobs) {
                float jm = 0.f;
                DH_HIP(hipEventElapsedTime(&jm, J->ev[0], J->ev[1]));
                J->ms = jm;
            }
            s->timed = true;
            s->planned = false;
        } else {
            if (!s->exec) {
                DH_HIP(hipStreamBeginCapture(s->origin, hipStreamCaptureModeThreadLocal));
                try {
                    DH_HIP(hipEventRecord(s->fork, s->origin));
                    enqueue_batch(s, false);
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(s->origin, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                DH_HIP(hipStreamEndCapture(s->origin, &s->graph));
                DH_HIP(hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
            }
            DH_HIP(hipGraphLaunch(s->exec, s->origin));
        }
        DH_HIP(hipEventRecord(s->done, s->origin));
        DH_HIP(hipStreamWaitEvent(caller, s->done, 0));
    });
}

extern "C" int dh_sched_info(const dh_sched *s, int32_t *n_cases, int32_t *n_jobs, double *last_ms,
                             int32_t *case_stream, double *job_ms)
{
    return dh_guarded([&] {
        DH_REQUIRE(s, "null scheduler");
        if (n_cases) *n_cases = (int32_t)s->cases.size();
        if (n_jobs) *n_jobs = (int32_t)s->jobs.size();
        if (last_ms) *last_ms = s->last_ms;
        if (case_stream)
            for (size_t c = 0; c < s->cases.size(); c++) case_stream[c] = s->cases[c].stream;
        if (job_ms)
            for (size_t j = 0; j < s->jobs.size(); j++) job_ms[j] = s->jobs[j]->ms;
    });
}
