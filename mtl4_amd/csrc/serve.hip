// serve.hip -- serving loop with continuous batching (no MTL4 counterpart).
//
// One host event loop replays the request trace in real time and drives the
// device:
//   arrivals  -> FIFO of pending requests
//   admission -> a request enters only when b, x and workspace for all its
//                right-hand sides fit the arena (device working-set cap); no
//                overtaking, so a large request is not starved
//   launch    -> whenever one of `streams` launch slots is free, every
//                admitted-and-waiting request (up to max_items right-hand
//                sides) is packed into one persistent batched launch on that
//                slot's stream; launches on different slots overlap
//   complete  -> an event per launch; its requests complete, memory returns
//                to the arena, waiting requests can be admitted
// Systems above big_rows rows go to `streams` worker threads that run the
// whole-GPU Krylov engine (matrix autotuned at set-up), each on a stream of
// its own, overlapping with each other and with the batched launches.
// Every request names its solver (itl::cg or itl::bicgstab with the
// identity, diagonal or ilu_0 preconditioner); a batched launch serves one
// solver.  Right-hand sides are generated on the device from a per-request
// seed.  The arena of request vectors is allocated when a trace is run: the
// smaller of the admission cap and the trace's total footprint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batched.hpp"
#include "kernels.hpp"
#include "krylov.hpp"
#include "precond.hpp"

namespace mtl4g {

// splitmix64 of (seed, i) mapped to [-1, 1)
__device__ __forceinline__ double hash_uniform(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

struct GenItem {
    int64_t b_off, x_off;
    int32_t n, pad;
    uint64_t seed;
};

// b = uniform(seed), x = 0 for every item (one workgroup per item)
__global__ void k_gen_items(const GenItem* __restrict__ items, int nitems, double* base) {
    for (int k = blockIdx.x; k < nitems; k += gridDim.x) {
        const GenItem g = items[k];
        for (int i = threadIdx.x; i < g.n; i += blockDim.x) {
            base[g.b_off + i] = hash_uniform(g.seed, (uint64_t)i);
            base[g.x_off + i] = 0.0;
        }
    }
}

// first-fit allocator over the device arena (units: doubles, 64-aligned)
struct Arena {
    int64_t cap = 0, used = 0, peak = 0;
    std::map<int64_t, int64_t> free_;
    void init(int64_t c) {
        cap = c;
        free_.clear();
        free_[0] = c;
        used = peak = 0;
    }
    static int64_t round(int64_t n) { return (n + 63) / 64 * 64; }
    bool alloc(int64_t n, int64_t& off) {
        n = round(n);
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                off = it->first;
                const int64_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                used += n;
                peak = std::max(peak, used);
                return true;
            }
        return false;
    }
    void release(int64_t off, int64_t n) {
        n = round(n);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Req {
    int32_t id, mat, nrhs, n;
    int32_t method = MTL4G_CG, pc = MTL4G_PC_IDENTITY;
    double arrival;
    uint64_t seed;
    int64_t off = 0, elems = 0, stride = 0;   // stride: b, x (and workspace) of one right-hand side
    bool big = false;
    int32_t solver() const { return MTL4G_SERVER_SOLVER(method, pc); }
};

struct Slot {
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<int32_t> reqs;
    int nitems = 0;
    mtl4g_batch_item* h_items = nullptr;
    GenItem* h_gen = nullptr;
    mtl4g_batch_result* h_res = nullptr;
    DevBuf<mtl4g_batch_item> d_items;
    DevBuf<GenItem> d_gen;
    DevBuf<mtl4g_batch_result> d_res;
    DevBuf<unsigned> d_queue;
};

}  // namespace mtl4g

struct mtl4g_server_s {
    mtl4g_batch_s* B = nullptr;
    mtl4g_server_params prm{};
    mtl4g::DevBuf<double> arena;
    std::vector<std::unique_ptr<mtl4g::Slot>> slots;
    // whole-GPU engine for large systems: matrix, diagonal and ilu_0
    std::vector<mtl4g_matrix_s*> big_A;
    std::vector<mtl4g_pc_s*> big_Pd, big_Pi;
    std::vector<hipStream_t> big_s;   // one per big-system worker
    ~mtl4g_server_s() {
        for (auto& slp : slots) {
            mtl4g::Slot& sl = *slp;
            if (sl.ev) (void)hipEventDestroy(sl.ev);
            if (sl.s) (void)hipStreamDestroy(sl.s);
            if (sl.h_items) (void)hipHostFree(sl.h_items);
            if (sl.h_gen) (void)hipHostFree(sl.h_gen);
            if (sl.h_res) (void)hipHostFree(sl.h_res);
        }
        for (auto* p : big_Pd) delete p;
        for (auto* p : big_Pi) delete p;
        for (auto* a : big_A) delete a;
        for (hipStream_t s : big_s) (void)hipStreamDestroy(s);
    }
};

namespace mtl4g {

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace mtl4g

using namespace mtl4g;

extern "C" {

int mtl4g_server_create(mtl4g_server* out, mtl4g_batch B, const mtl4g_server_params* prm) {
    if (!out || !B || !prm || prm->streams < 1 || prm->streams > 16 || prm->max_items < 1 || prm->cap_bytes <= 0 ||
        (prm->pc_kind != MTL4G_PC_IDENTITY && prm->pc_kind != MTL4G_PC_DIAGONAL && prm->pc_kind != MTL4G_PC_ILU0))
        return MTL4G_ERR_INVALID;
    *out = nullptr;
    mtl4g_server_s* S_ = new (std::nothrow) mtl4g_server_s;
    if (!S_) return MTL4G_ERR_OOM;
    auto fail = [&](int st) { delete S_; return st; };
    S_->B = B;
    S_->prm = *prm;
    int st = batch_ensure_values(B, prm->precision);
    if (st) return fail(st);
    if ((st = batch_ensure_ilu0(B))) return fail(st);   // ilu_0 requests on the batched kernel
    for (int k = 0; k < prm->streams; k++) S_->slots.emplace_back(new Slot);
    for (auto& slp : S_->slots) {
        Slot& sl = *slp;
        if (hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking) || hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) ||
            hipHostMalloc((void**)&sl.h_items, sizeof(mtl4g_batch_item) * prm->max_items, hipHostMallocDefault) ||
            hipHostMalloc((void**)&sl.h_gen, sizeof(GenItem) * prm->max_items, hipHostMallocDefault) ||
            hipHostMalloc((void**)&sl.h_res, sizeof(mtl4g_batch_result) * prm->max_items, hipHostMallocDefault))
            return fail(hip_status(hipErrorOutOfMemory, "server slot", __FILE__, __LINE__));
        if ((st = sl.d_items.alloc(prm->max_items)) || (st = sl.d_gen.alloc(prm->max_items)) ||
            (st = sl.d_res.alloc(prm->max_items)) || (st = sl.d_queue.alloc(1)))
            return fail(st);
    }
    // large systems: their own CSR handle and preconditioner for the whole-GPU engine
    S_->big_A.assign(B->nsys, nullptr);
    S_->big_Pd.assign(B->nsys, nullptr);
    S_->big_Pi.assign(B->nsys, nullptr);
    for (int32_t k = 0; k < B->nsys; k++) {
        const int32_t a = B->h_sys_row[k], e = B->h_sys_row[k + 1], n = e - a;
        if (n <= prm->big_rows) continue;
        const int32_t j0 = B->h_rowptr[a], j1 = B->h_rowptr[e];
        std::vector<int32_t> rp(n + 1), ci(j1 - j0);
        for (int32_t r = 0; r <= n; r++) rp[r] = B->h_rowptr[a + r] - j0;
        for (int32_t j = j0; j < j1; j++) ci[j - j0] = B->h_col[j] - a;
        mtl4g_matrix A = nullptr;
        if ((st = mtl4g_csr_create(&A, n, n, j1 - j0, rp.data(), ci.data(), B->h_val.data() + j0, 0))) return fail(st);
        S_->big_A[k] = A;
        // kernel and low-precision copies settled here: the workers only read the matrix
        if ((st = mtl4g_csr_autotune(A, nullptr, nullptr))) return fail(st);
        if (prm->precision != MTL4G_F64 && (st = ensure_lowp_values(A, prm->precision, nullptr))) return fail(st);
        mtl4g_pc P = nullptr;
        const int fmt = prm->precision == MTL4G_F64 ? MTL4G_F64 : MTL4G_F32;
        if ((st = mtl4g_pc_create(&P, A, MTL4G_PC_DIAGONAL, fmt))) return fail(st);
        S_->big_Pd[k] = P;
        if ((st = mtl4g_pc_create(&P, A, MTL4G_PC_ILU0, fmt))) return fail(st);
        S_->big_Pi[k] = P;
    }
    S_->big_s.assign(prm->streams, nullptr);
    for (hipStream_t& bs : S_->big_s)
        if (hipStreamCreateWithFlags(&bs, hipStreamNonBlocking)) return fail(MTL4G_ERR_HIP);
    *out = S_;
    return MTL4G_OK;
}

int mtl4g_server_destroy(mtl4g_server S_) {
    delete S_;
    return MTL4G_OK;
}

int mtl4g_server_run(mtl4g_server S_, int32_t nreq, const int32_t* mat, const int32_t* nrhs,
                     const int32_t* solver, const double* arrival_s, const uint64_t* seed, double* latency_s,
                     int32_t* iters, double* x_out, const int64_t* x_off, mtl4g_server_stats* stats) {
    if (!S_ || nreq < 0 || (nreq && (!mat || !nrhs || !arrival_s || !seed || !latency_s))) return MTL4G_ERR_INVALID;
    mtl4g_batch_s* B = S_->B;
    const mtl4g_server_params& prm = S_->prm;
    const int64_t cap_elems = (int64_t)(prm.cap_bytes / 8.0);
    std::vector<Req> reqs(nreq);
    std::vector<int64_t> iter_off(nreq + 1, 0);
    for (int32_t r = 0; r < nreq; r++) {
        if (mat[r] < 0 || mat[r] >= B->nsys || nrhs[r] < 1 || nrhs[r] > prm.max_items) {
            set_error("request %d: matrix %d / nrhs %d out of range", r, mat[r], nrhs[r]);
            return MTL4G_ERR_INVALID;
        }
        Req& q = reqs[r];
        q.id = r;
        q.mat = mat[r];
        q.nrhs = nrhs[r];
        q.n = B->h_sys_row[q.mat + 1] - B->h_sys_row[q.mat];
        q.arrival = arrival_s[r];
        q.seed = seed[r];
        q.big = S_->big_A[q.mat] != nullptr;
        if (solver) {
            q.method = solver[r] >> 8;
            q.pc = solver[r] & 0xff;
        } else {
            q.method = MTL4G_CG;
            q.pc = prm.pc_kind;
        }
        if ((q.method != MTL4G_CG && q.method != MTL4G_BICGSTAB) ||
            (q.pc != MTL4G_PC_IDENTITY && q.pc != MTL4G_PC_DIAGONAL && q.pc != MTL4G_PC_ILU0)) {
            set_error("request %d: solver code %d (method %d, preconditioner %d) not served", r,
                      solver ? solver[r] : -1, q.method, q.pc);
            return MTL4G_ERR_INVALID;
        }
        q.stride = Arena::round(q.n) * 2 +
                   (q.big ? 0 : Arena::round(batch_ws_elems(q.n, MTL4G_F64, q.method, q.pc)));
        q.elems = q.stride * q.nrhs;
        if (q.elems > cap_elems) {
            set_error("request %d needs %lld bytes, above the cap", r, (long long)(q.elems * 8));
            return MTL4G_ERR_INVALID;
        }
        iter_off[r + 1] = iter_off[r] + q.nrhs;
    }
    // the arena: admission never holds more than the cap, nor more than the
    // whole trace at once
    {
        int64_t total = 0;
        for (const Req& q : reqs) total += Arena::round(q.elems);
        const int64_t want = std::max<int64_t>(64, std::min(cap_elems, total));
        if ((int64_t)S_->arena.n < want) {
            S_->arena.free();
            M4_TRY(S_->arena.alloc((size_t)want));
        }
    }
    Arena arena;
    arena.init(std::min<int64_t>((int64_t)S_->arena.n, cap_elems));
    double* base = S_->arena.p;
    std::vector<int32_t> order(nreq);
    for (int32_t r = 0; r < nreq; r++) order[r] = r;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return reqs[a].arrival < reqs[b].arrival; });

    // big-system workers
    std::mutex mu;
    std::condition_variable cv;
    // whole-GPU work is handed out one right-hand side at a time, so the
    // right-hand sides of one request run on several workers at once
    std::deque<std::pair<int32_t, int32_t>> big_todo;   // (request, right-hand side)
    std::vector<int32_t> big_left(reqs.size(), 0);       // right-hand sides still running
    std::deque<std::pair<int32_t, double>> big_done;
    std::atomic<bool> stop{false};
    std::atomic<int> worker_err{0};
    std::atomic<int32_t> big_errors{0}, big_solves{0};
    auto worker_main = [&](hipStream_t bs) {
        (void)hipSetDevice(B->device);
        Workspace<double> W64;
        Workspace<float> W32;
        DevBuf<float> x32;
        DevBuf<GenItem> dg;
        if (dg.alloc(1)) {
            worker_err = MTL4G_ERR_OOM;
            return;
        }
        for (;;) {
            int32_t id, k;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop.load() || !big_todo.empty(); });
                if (big_todo.empty()) return;
                id = big_todo.front().first;
                k = big_todo.front().second;
                big_todo.pop_front();
            }
            const Req& q = reqs[id];
            const int64_t pn = Arena::round(q.n);
            mtl4g_pc_s* P = q.pc == MTL4G_PC_DIAGONAL ? S_->big_Pd[q.mat]
                            : q.pc == MTL4G_PC_ILU0    ? S_->big_Pi[q.mat]
                                                       : nullptr;
            do {
                GenItem g{q.off + k * q.stride, q.off + k * q.stride + pn, q.n, 0, q.seed * 131 + (uint64_t)k};
                if (hipMemcpyAsync(dg.p, &g, sizeof g, hipMemcpyHostToDevice, bs)) {
                    worker_err = MTL4G_ERR_HIP;
                    break;
                }
                hipLaunchKernelGGL(k_gen_items, dim3(1), dim3(256), 0, bs, dg.p, 1, base);
                double* bb = base + g.b_off;
                double* xx = base + g.x_off;
                mtl4g_iteration it{};
                it.max_iter = prm.max_iter;
                it.rtol = prm.rtol;
                // norm_r0 = ||b|| as basic_iteration(b, ...)
                double nb = 0;
                if (mtl4g_two_norm(q.n, bb, &nb, bs)) { worker_err = MTL4G_ERR_HIP; break; }
                it.norm_r0 = nb;
                int st;
                if (prm.precision == MTL4G_F64) {
                    st = krylov_run<double>(S_->big_A[q.mat], P, q.method, MTL4G_F64, W64, xx, bb, &it, nullptr, 0, bs);
                } else {
                    if (q.method == MTL4G_CG && x32.n < (size_t)q.n && x32.alloc(q.n)) {
                        worker_err = MTL4G_ERR_OOM;
                        break;
                    }
                    int32_t rf = 0;
                    st = mixed_run(S_->big_A[q.mat], P, q.method, prm.precision, 0.1, 0, W32, W64, x32.p, xx, bb, &it,
                                   &rf, bs);
                }
                if (st != 0) big_errors++;
                big_solves++;
                if (iters) iters[iter_off[id] + k] = it.i;
            } while (false);
            (void)hipStreamSynchronize(bs);
            std::lock_guard<std::mutex> lk(mu);
            if (--big_left[id] == 0) big_done.emplace_back(id, now_s());
        }
    };
    std::vector<std::thread> workers;
    for (hipStream_t bs : S_->big_s) workers.emplace_back(worker_main, bs);

    const double t0 = now_s();
    size_t next = 0;
    std::deque<int32_t> pending;   // arrived, not yet admitted (FIFO)
    std::deque<int32_t> admitted;  // admitted small requests waiting for a launch slot
    int32_t completed = 0, launches = 0, max_inflight = 0, errors = 0;
    int64_t solves = 0;
    double t_last = t0;
    int status = MTL4G_OK;
    auto complete = [&](int32_t id, double t) {
        latency_s[id] = t - (t0 + reqs[id].arrival);
        t_last = std::max(t_last, t);
        if (x_out && x_off) {
            const Req& q = reqs[id];
            const int64_t pn = Arena::round(q.n);
            for (int32_t k = 0; k < q.nrhs; k++)
                (void)hipMemcpy(x_out + x_off[id] + (int64_t)k * q.n, base + q.off + k * q.stride + pn,
                                (size_t)q.n * 8, hipMemcpyDeviceToHost);
        }
        arena.release(reqs[id].off, reqs[id].elems);
        completed++;
    };
    while (completed < nreq && status == MTL4G_OK) {
        bool progress = false;
        const double now = now_s();
        // arrivals
        while (next < order.size() && t0 + reqs[order[next]].arrival <= now) {
            pending.push_back(order[next++]);
            progress = true;
        }
        // completions of batched launches
        int inflight = 0;
        for (auto& slp : S_->slots) {
            Slot& sl = *slp;
            if (!sl.busy) continue;
            hipError_t q = hipEventQuery(sl.ev);
            if (q == hipErrorNotReady) {
                inflight++;
                continue;
            }
            if (q != hipSuccess) {
                status = hip_status(q, "batched launch", __FILE__, __LINE__);
                break;
            }
            const double t = now_s();
            int item = 0;
            for (int32_t id : sl.reqs) {
                for (int32_t k = 0; k < reqs[id].nrhs; k++, item++) {
                    const mtl4g_batch_result& res = sl.h_res[item];
                    if (res.error || res.status) errors++;
                    if (iters) iters[iter_off[id] + k] = res.iterations;
                    solves++;
                }
                complete(id, t);
            }
            sl.busy = false;
            sl.reqs.clear();
            progress = true;
        }
        // completions of the big-system worker
        {
            std::lock_guard<std::mutex> lk(mu);
            while (!big_done.empty()) {
                auto d = big_done.front();
                big_done.pop_front();
                solves += reqs[d.first].nrhs;
                complete(d.first, d.second);
                progress = true;
            }
        }
        if (worker_err) status = worker_err;
        // admission: FIFO while the working set fits
        while (!pending.empty()) {
            Req& q = reqs[pending.front()];
            int64_t off;
            if (!arena.alloc(q.elems, off)) break;
            q.off = off;
            pending.pop_front();
            if (q.big) {
                std::lock_guard<std::mutex> lk(mu);
                big_left[q.id] = q.nrhs;
                for (int32_t k = 0; k < q.nrhs; k++) big_todo.emplace_back(q.id, k);
                cv.notify_all();
            } else {
                admitted.push_back(q.id);
            }
            progress = true;
        }
        // launches: every free slot takes what is admitted
        for (auto& slp : S_->slots) {
            Slot& sl = *slp;
            if (sl.busy || admitted.empty()) continue;
            // one solver and one size class per launch: the oldest admitted
            // request's, then every admitted request of that solver and class
            // in order (the others keep their place).  The class sets the
            // workgroup size (one workgroup runs one system's whole loop);
            // the workgroup's reductions depend on it, so it depends on the
            // system alone: a request's result is the same alone and in a mix.
            auto size_class = [](int32_t n) { return n > 8192 ? 2 : n > 2048 ? 1 : 0; };
            const int32_t solv = reqs[admitted.front()].solver();
            const int cls = size_class(reqs[admitted.front()].n);
            const int method = solv >> 8, pc_kind = solv & 0xff;
            int nit = 0;
            for (auto ita = admitted.begin(); ita != admitted.end();) {
                const Req& q = reqs[*ita];
                if (q.solver() != solv || size_class(q.n) != cls) { ++ita; continue; }
                if (nit + q.nrhs > prm.max_items || (prm.per_request && nit > 0)) break;
                const int64_t pn = Arena::round(q.n);
                for (int32_t k = 0; k < q.nrhs; k++, nit++) {
                    const int64_t o = q.off + k * q.stride;
                    sl.h_gen[nit] = GenItem{o, o + pn, q.n, 0, q.seed * 131 + (uint64_t)k};
                    sl.h_items[nit] = mtl4g_batch_item{q.mat, prm.max_iter, o, o + pn, o + 2 * pn, prm.rtol, 0.0};
                }
                sl.reqs.push_back(q.id);
                ita = admitted.erase(ita);
            }
            sl.nitems = nit;
            hipStream_t s = sl.s;
            int st = 0;
            if (hipMemcpyAsync(sl.d_gen.p, sl.h_gen, nit * sizeof(GenItem), hipMemcpyHostToDevice, s) ||
                hipMemcpyAsync(sl.d_items.p, sl.h_items, nit * sizeof(mtl4g_batch_item), hipMemcpyHostToDevice, s))
                st = MTL4G_ERR_HIP;
            if (!st) {
                hipLaunchKernelGGL(k_gen_items, dim3(std::min(nit, 4096)), dim3(256), 0, s, sl.d_gen.p, nit, base);
                BatchShape sh = B->shape[prm.precision == MTL4G_F64 ? 0 : prm.precision == MTL4G_F32 ? 1 : 2];
                if (cls == 2) sh = BatchShape{1024, prm.precision == MTL4G_F64 ? 1 : 2};
                else if (cls == 1) sh = BatchShape{512, 2};
                st = batch_launch(B, method, pc_kind, prm.precision, sh, sl.d_items.p, nit, base, base, base,
                                  sl.d_res.p, sl.d_queue.p, s);
            }
            if (!st && (hipMemcpyAsync(sl.h_res, sl.d_res.p, nit * sizeof(mtl4g_batch_result), hipMemcpyDeviceToHost, s) ||
                        hipEventRecord(sl.ev, s)))
                st = MTL4G_ERR_HIP;
            if (st) {
                status = st;
                break;
            }
            sl.busy = true;
            launches++;
            progress = true;
        }
        int busy = 0;
        for (auto& slp : S_->slots) busy += slp->busy;
        max_inflight = std::max(max_inflight, busy);
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        stop = true;
        cv.notify_all();
    }
    for (std::thread& w : workers) w.join();
    for (auto& slp : S_->slots) {
        if (slp->busy) (void)hipStreamSynchronize(slp->s);
        slp->busy = false;
        slp->reqs.clear();
    }
    if (stats) {
        stats->wall_s = t_last - t0;
        stats->solves = solves;
        stats->launches = launches;
        stats->max_inflight = max_inflight;
        stats->peak_bytes = (double)arena.peak * 8.0;
        stats->errors = errors + big_errors.load();
        stats->big_solves = big_solves.load();
    }
    return status;
}

}  // extern "C"
