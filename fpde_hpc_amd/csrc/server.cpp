// Continuous-batching request server on one GPU.
//
// Requests (1D / 3D / distributed-order solves of uneven size) queue in
// arrival order.  A scheduling round admits queued requests while the
// device-memory budget (projected history bytes, fpde_solver_footprint) and
// the active-slot limit allow, advances every active solve by a quantum of
// steps, and retires solves whose streams have drained, freeing their memory
// and slot for the next requests.  The 1D solves of the cooperative block
// form advance together: as many as fit on the GPU at once share one
// cooperative launch per history block (fpde_solver_advance_batch).  The
// others (3D, per-iteration 1D) advance on their own HIP streams, so their
// work overlaps on the device.  State stays on the device for the whole life
// of a request; only u0 and the final field cross PCIe.
#include "fpde_internal.h"

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <vector>

using namespace fpde;

// ---------------------------------------------------------------------------
// DeviceArena (declared in fpde_internal.h)
DeviceArena::~DeviceArena() {
    for (auto& sg : segs_) hipFree(sg.first);
}

void* DeviceArena::alloc(size_t bytes) {
    const size_t need = (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
    for (auto it = free_.begin(); it != free_.end(); ++it) {       // first fit
        if (it->second.size < need) continue;
        char* p = it->first;
        Block b = it->second;
        free_.erase(it);
        if (b.size > need) free_[p + need] = Block{b.size - need, b.seg};
        live_[p] = Block{need, b.seg};
        in_use_ += need;
        return p;
    }
    const size_t seg = std::max(seg_bytes_, need);
    if (reserved_ + seg > limit_) return nullptr;
    void* base = nullptr;
    if (hipMalloc(&base, seg) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
    const int id = (int)segs_.size();
    segs_.push_back({(char*)base, seg});
    reserved_ += seg;
    if (seg > need) free_[(char*)base + need] = Block{seg - need, id};
    live_[(char*)base] = Block{need, id};
    in_use_ += need;
    return base;
}

void DeviceArena::release(void* ptr) {
    auto it = live_.find((char*)ptr);
    if (it == live_.end()) return;
    char* p = it->first;
    Block b = it->second;
    live_.erase(it);
    in_use_ -= b.size;
    // coalesce with the free neighbours of the same segment
    auto nx = free_.lower_bound(p);
    if (nx != free_.end() && nx->first == p + b.size && nx->second.seg == b.seg) {
        b.size += nx->second.size;
        nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second.size == p && pv->second.seg == b.seg) {
            pv->second.size += b.size;
            return;
        }
    }
    free_[p] = b;
}

namespace {
struct Request {
    int64_t id = 0;
    fpde_spec spec{};
    std::vector<double> u0;
    std::vector<double> result;
    fpde_solver* solver = nullptr;
    size_t bytes = 0;
    double submit_s = 0, start_s = 0, finish_s = 0;
    int status = -1;       // -1 queued/active, 0 done, 1 rejected, 2 cancelled, 3 failed
    bool active = false;
};
}  // namespace

struct fpde_server {
    size_t budget = 0;
    size_t used = 0;
    int max_active = 8;
    // device memory of every admitted solve (segments reused across requests)
    // and one stream per active slot, created once
    std::unique_ptr<DeviceArena> arena;
    std::vector<hipStream_t> streams;
    std::vector<hipStream_t> free_streams;
    int64_t next_id = 1;
    int64_t finished = 0;
    std::deque<Request*> queue;
    std::vector<Request*> active;
    std::map<Request*, hipStream_t> slot_stream;
    std::map<int64_t, std::unique_ptr<Request>> all;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

int64_t n_unknowns(const fpde_spec& s) {
    return s.dims == 1 ? (int64_t)s.nodes : (int64_t)s.nodes * s.nodes * s.nodes;
}

void retire(fpde_server* srv, Request* r, int status) {
    if (r->solver) {
        if (status == 0) {
            r->result.resize((size_t)n_unknowns(r->spec));
            if (fpde_solver_get_solution(r->solver, r->result.data()) != FPDE_OK) status = 3;
        }
        fpde_solver_destroy(r->solver);      // synchronises the slot's stream
        r->solver = nullptr;
        srv->used -= r->bytes;
    }
    auto st = srv->slot_stream.find(r);
    if (st != srv->slot_stream.end()) {
        srv->free_streams.push_back(st->second);
        srv->slot_stream.erase(st);
    }
    r->status = status;
    r->finish_s = srv->now();
    r->active = false;
    r->u0.clear();
    r->u0.shrink_to_fit();
    srv->finished++;
}

// admit from the head of the queue while memory and slots allow
fpde_status admit(fpde_server* srv) {
    while (!srv->queue.empty() && (int)srv->active.size() < srv->max_active) {
        Request* r = srv->queue.front();
        if (r->status != -1) { srv->queue.pop_front(); continue; }
        if (srv->used + r->bytes > srv->budget) break;
        if (srv->free_streams.empty()) break;
        hipStream_t st = srv->free_streams.back();
        fpde_status rc = solver_create_in(&r->spec, srv->arena.get(), st, &r->solver);
        if (rc == FPDE_ERR_NOMEM) { r->solver = nullptr; break; }   // retry when memory frees up
        srv->queue.pop_front();
        if (rc != FPDE_OK) { r->status = 1; r->finish_s = srv->now(); srv->finished++; continue; }
        srv->free_streams.pop_back();                                // the slot's stream
        srv->slot_stream[r] = st;
        srv->used += r->bytes;
        rc = fpde_solver_set_initial(r->solver, r->u0.data());
        if (rc != FPDE_OK) { retire(srv, r, 3); continue; }
        r->start_s = srv->now();
        r->active = true;
        srv->active.push_back(r);
    }
    return FPDE_OK;
}

}  // namespace

extern "C" {

fpde_status fpde_server_create(size_t mem_budget_bytes, int max_active, fpde_server** out) {
    if (!out || max_active < 1 || max_active > 1024) { set_error("bad arguments"); return FPDE_ERR_ARG; }
    auto s = std::make_unique<fpde_server>();
    s->budget = mem_budget_bytes;
    s->max_active = max_active;
    // segments of at least 8 GiB (or an eighth of the budget), up to the
    // budget plus one segment of slack for what first-fit leaves between blocks
    const size_t seg = std::max<size_t>(std::min<size_t>(mem_budget_bytes / 8 + 1, (size_t)8 << 30), (size_t)64 << 20);
    s->arena = std::make_unique<DeviceArena>(seg, mem_budget_bytes + seg);
    for (int i = 0; i < max_active; i++) {
        hipStream_t st = nullptr;
        hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (hipStream_t x : s->streams) hipStreamDestroy(x);
            set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
            return FPDE_ERR_HIP;
        }
        s->streams.push_back(st);
    }
    s->free_streams = s->streams;
    *out = s.release();
    return FPDE_OK;
}

fpde_status fpde_server_destroy(fpde_server* srv) {
    if (!srv) return FPDE_OK;
    for (auto& kv : srv->all)
        if (kv.second->solver) fpde_solver_destroy(kv.second->solver);
    for (hipStream_t st : srv->streams) hipStreamDestroy(st);
    delete srv;                          // the arena frees its segments
    return FPDE_OK;
}

fpde_status fpde_server_submit(fpde_server* srv, const fpde_spec* spec, const double* u0, int64_t* id_out) {
    if (!srv || !spec || !u0) { set_error("null argument"); return FPDE_ERR_ARG; }
    auto r = std::make_unique<Request>();
    r->id = srv->next_id++;
    r->spec = *spec;
    r->submit_s = srv->now();
    r->bytes = fpde_solver_footprint(spec);
    if (id_out) *id_out = r->id;
    if (r->bytes == 0 || r->bytes > srv->budget) {
        r->status = 1;
        r->finish_s = r->submit_s;
        srv->finished++;
    } else {
        r->u0.assign(u0, u0 + n_unknowns(*spec));
        srv->queue.push_back(r.get());
    }
    srv->all[r->id] = std::move(r);
    return FPDE_OK;
}

fpde_status fpde_server_cancel(fpde_server* srv, int64_t id) {
    if (!srv) { set_error("null server"); return FPDE_ERR_ARG; }
    auto it = srv->all.find(id);
    if (it == srv->all.end()) { set_error("unknown request"); return FPDE_ERR_ARG; }
    Request* r = it->second.get();
    if (r->status != -1) return FPDE_OK;
    if (r->active) {
        for (size_t i = 0; i < srv->active.size(); i++)
            if (srv->active[i] == r) { srv->active.erase(srv->active.begin() + i); break; }
        retire(srv, r, 2);
    } else {
        r->status = 2;
        r->finish_s = srv->now();
        srv->finished++;
    }
    return FPDE_OK;
}

int fpde_server_poll(fpde_server* srv, int quantum) {
    if (!srv) return -1;
    admit(srv);
    std::vector<Request*> coop, riders;
    for (Request* r : srv->active) {
        if (fpde_solver_step(r->solver) >= r->spec.steps) continue;
        if (solver_batchable(r->solver)) { coop.push_back(r); continue; }
        if (solver_3d_batchable(r->solver)) { riders.push_back(r); continue; }
        if (fpde_solver_advance(r->solver, quantum) != FPDE_OK) r->status = 3;
    }
    // the 3D requests in batched LOD steps (groups of kMax3DBatch)
    for (size_t g0 = 0; g0 < riders.size(); g0 += kMax3DBatch) {
        std::vector<fpde_solver*> grp;
        for (size_t i = g0; i < riders.size() && i < g0 + kMax3DBatch; i++) grp.push_back(riders[i]->solver);
        if (advance_3d_batch(grp.data(), (int)grp.size(), quantum) != FPDE_OK)
            for (size_t i = g0; i < riders.size() && i < g0 + kMax3DBatch; i++) riders[i]->status = 3;
    }
    // the cooperative 1D solves, in groups whose chunk workgroups fit on the
    // GPU together: one batched launch per history block for each group (the
    // call returns when the group's quantum is done; the other requests'
    // work, enqueued above, runs meanwhile)
    size_t g0 = 0;
    while (g0 < coop.size()) {
        std::vector<fpde_solver*> grp;
        int64_t wgs = 0;
        size_t g1 = g0;
        while (g1 < coop.size() && (int)grp.size() < kMaxBatch &&
               (grp.empty() || step1d_batch_supported(wgs + solver_chunks(coop[g1]->solver)))) {
            wgs += solver_chunks(coop[g1]->solver);
            grp.push_back(coop[g1]->solver);
            g1++;
        }
        if (fpde_solver_advance_batch(grp.data(), (int)grp.size(), quantum) != FPDE_OK) {
            // a solve whose barrier timed out or whose launch failed reports
            // it on sync; the others of the group go on next round
            bool any = false;
            for (size_t i = g0; i < g1; i++)
                if (fpde_solver_sync(coop[i]->solver) != FPDE_OK) { coop[i]->status = 3; any = true; }
            if (!any)
                for (size_t i = g0; i < g1; i++) coop[i]->status = 3;
        }
        g0 = g1;
    }
    // retire drained solves
    std::vector<Request*> keep;
    for (Request* r : srv->active) {
        bool done = fpde_solver_step(r->solver) >= r->spec.steps && fpde_solver_query(r->solver) != 0;
        if (r->status == 3) { retire(srv, r, 3); continue; }
        if (done) { retire(srv, r, 0); continue; }
        keep.push_back(r);
    }
    srv->active.swap(keep);
    admit(srv);
    return (int)(srv->active.size() + srv->queue.size());
}

fpde_status fpde_server_run(fpde_server* srv, int64_t stop_after) {
    if (!srv) { set_error("null server"); return FPDE_ERR_ARG; }
    while (true) {
        int left = fpde_server_poll(srv, 256);   // four 64-step blocks per round (measured best)
        if (left < 0) return FPDE_ERR_STATE;
        if (left == 0) break;
        if (stop_after > 0 && srv->finished >= stop_after) break;
    }
    return FPDE_OK;
}

fpde_status fpde_server_result(fpde_server* srv, int64_t id, fpde_result_info* info, double* u_out, size_t u_len) {
    if (!srv) { set_error("null server"); return FPDE_ERR_ARG; }
    auto it = srv->all.find(id);
    if (it == srv->all.end()) { set_error("unknown request"); return FPDE_ERR_ARG; }
    Request* r = it->second.get();
    if (info) {
        info->id = r->id;
        info->submit_s = r->submit_s;
        info->start_s = r->start_s;
        info->finish_s = r->finish_s;
        info->status = r->status;
        info->steps_done = r->solver ? fpde_solver_step(r->solver) : (r->status == 0 ? r->spec.steps : 0);
    }
    if (u_out) {
        if (r->status != 0) { set_error("request has no result"); return FPDE_ERR_STATE; }
        if (u_len < r->result.size()) { set_error("output too small"); return FPDE_ERR_ARG; }
        std::copy(r->result.begin(), r->result.end(), u_out);
    }
    return FPDE_OK;
}

double fpde_server_clock(fpde_server* srv) { return srv ? srv->now() : 0.0; }

}  // extern "C"
