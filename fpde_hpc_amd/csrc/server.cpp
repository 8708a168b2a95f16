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

#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <vector>

using namespace fpde;

namespace {
struct Request {
    int64_t id = 0;
    fpde_spec spec{};
    std::vector<double> u0;
    std::vector<double> result;
    fpde_solver* solver = nullptr;
    size_t bytes = 0;
    double submit_s = 0, start_s = 0, finish_s = 0;
    int status = -1;       // -1 queued/active, 0 done, 1 rejected, 2 cancelled
    bool active = false;
};
}  // namespace

struct fpde_server {
    size_t budget = 0;
    size_t used = 0;
    int max_active = 8;
    int64_t next_id = 1;
    int64_t finished = 0;
    std::deque<Request*> queue;
    std::vector<Request*> active;
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
            if (fpde_solver_get_solution(r->solver, r->result.data()) != FPDE_OK) status = 2;
        }
        fpde_solver_destroy(r->solver);
        r->solver = nullptr;
        srv->used -= r->bytes;
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
        fpde_status rc = fpde_solver_create(&r->spec, &r->solver);
        if (rc == FPDE_ERR_NOMEM) { r->solver = nullptr; break; }   // retry when memory frees up
        srv->queue.pop_front();
        if (rc != FPDE_OK) { r->status = 1; r->finish_s = srv->now(); srv->finished++; continue; }
        srv->used += r->bytes;
        rc = fpde_solver_set_initial(r->solver, r->u0.data());
        if (rc != FPDE_OK) { retire(srv, r, 2); continue; }
        r->start_s = srv->now();
        r->active = true;
        srv->active.push_back(r);
    }
    return FPDE_OK;
}

}  // namespace

extern "C" {

fpde_status fpde_server_create(size_t mem_budget_bytes, int max_active, fpde_server** out) {
    if (!out || max_active < 1) { set_error("bad arguments"); return FPDE_ERR_ARG; }
    auto* s = new fpde_server;
    s->budget = mem_budget_bytes;
    s->max_active = max_active;
    *out = s;
    return FPDE_OK;
}

fpde_status fpde_server_destroy(fpde_server* srv) {
    if (!srv) return FPDE_OK;
    for (auto& kv : srv->all)
        if (kv.second->solver) fpde_solver_destroy(kv.second->solver);
    delete srv;
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
    std::vector<Request*> coop;
    for (Request* r : srv->active) {
        if (fpde_solver_step(r->solver) >= r->spec.steps) continue;
        if (solver_batchable(r->solver)) { coop.push_back(r); continue; }
        if (fpde_solver_advance(r->solver, quantum) != FPDE_OK) r->status = 2;
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
        if (fpde_solver_advance_batch(grp.data(), (int)grp.size(), quantum) != FPDE_OK)
            for (size_t i = g0; i < g1; i++) coop[i]->status = 2;
        g0 = g1;
    }
    // retire drained solves
    std::vector<Request*> keep;
    for (Request* r : srv->active) {
        bool done = fpde_solver_step(r->solver) >= r->spec.steps && fpde_solver_query(r->solver) != 0;
        if (r->status == 2) { retire(srv, r, 2); continue; }
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
