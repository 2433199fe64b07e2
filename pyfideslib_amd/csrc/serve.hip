// Request loop with continuous batching over the batched key-switch plans.
//
// FIDESlib has no serving layer; this is the engine's own.  Jobs arrive with
// their inputs resident in HBM (ciphertexts, plaintexts, relinearization and
// rotation keys stay loaded) and are admitted while a slot is free and their
// device-memory footprint fits the budget.  Each step (tick) advances every
// active job by one stage: the jobs' local ops (ct x pt, scalar ops, adds,
// rescales) are enqueued, then the key-switched ops of ALL active jobs --
// HMult+relin+rescale and rotations, at whatever levels -- run as one batch
// (one launch per stage), then the jobs' follow-up local ops.  Jobs that
// arrive while others are in flight join at the next step, so requests of
// different kinds and levels share launches.  All work is ordered on the
// context stream; a job is complete when the event recorded after its last
// op has fired.
//
// Job kinds (the north-star serving mix):
//   MULT    out = a * b, relinearized and rescaled
//   ROTATE  out = rotate(a, index)
//   LOGREG  encrypted logistic-regression scoring: y = x * w (plaintext
//           weights, rescaled), rotate-and-sum over n_rot power-of-two
//           rotations, then sigma(y) ~ c0 + c1 y + c3 y^3.
#include <algorithm>
#include <deque>
#include <map>
#include <vector>

#include "context.hpp"

namespace fides {

namespace {

enum JobKind { JOB_MULT = 0, JOB_ROTATE = 1, JOB_LOGREG = 2 };

struct Job {
    int64_t id = 0;
    int kind = JOB_MULT;
    const Ciphertext* a = nullptr;
    const Ciphertext* b = nullptr;
    const Plaintext* w = nullptr;
    Ciphertext* out = nullptr;
    int index = 0, n_rot = 0;
    double c0 = 0, c1 = 0, c3 = 0;
    size_t bytes = 0;  // admitted device footprint
    int step = 0;
    std::vector<std::unique_ptr<Ciphertext>> tmp;  // y, t, y2, u, y3, v (LOGREG)
};

}  // namespace

class Server {
   public:
    Server(Context& c, int max_slots, uint64_t budget) : c_(c), max_slots_(max_slots), budget_(budget) {}
    ~Server() {
        (void)hipStreamSynchronize(c_.stream);
        for (auto& e : events_) (void)hipEventDestroy(e.ev);
    }

    int64_t submit(std::unique_ptr<Job> j) {
        j->id = next_id_++;
        j->bytes = footprint(*j);
        if (j->bytes > budget_) throw Error(E_OOM, "job footprint exceeds the server's device-memory budget");
        const int64_t id = j->id;
        queue_.push_back(std::move(j));
        return id;
    }

    // One scheduling step.  Returns the number of jobs still active or queued.
    int step() {
        admit();
        if (active_.empty()) return (int)queue_.size();
        std::vector<KsItem> items;
        for (auto& j : active_) pre(*j, items);
        ks_batch(c_, items);
        Completion done;
        for (auto it = active_.begin(); it != active_.end();) {
            Job& j = **it;
            if (post(j)) {
                used_ -= j.bytes;
                done.jobs.push_back(std::move(*it));
                it = active_.erase(it);
            } else {
                ++it;
            }
        }
        if (!done.jobs.empty()) {
            hip_check(hipEventCreateWithFlags(&done.ev, hipEventDisableTiming), "hipEventCreate");
            hip_check(hipEventRecord(done.ev, c_.stream), "hipEventRecord");
            events_.push_back(std::move(done));
        }
        return (int)(active_.size() + queue_.size());
    }

    // Ids of jobs whose device work has completed (each reported once); their
    // temporaries go back to the pool.
    std::vector<int64_t> poll() {
        std::vector<int64_t> ids;
        while (!events_.empty()) {
            hipError_t q = hipEventQuery(events_.front().ev);
            if (q == hipErrorNotReady) break;
            hip_check(q, "hipEventQuery");
            (void)hipEventDestroy(events_.front().ev);
            for (auto& j : events_.front().jobs) ids.push_back(j->id);
            events_.pop_front();
        }
        return ids;
    }

    int active() const { return (int)active_.size(); }
    int queued() const { return (int)queue_.size(); }
    uint64_t used_bytes() const { return used_; }

   private:
    size_t ct_bytes() const { return c_.ct_words() * 8; }
    // Scratch of one key-switched op at `level` in a batch plan (batch.hip layout).
    size_t ks_bytes(int level, bool mult) const {
        const size_t nl = level + 1, npos = nl + c_.K, nd = c_.num_digits(level);
        return ((mult ? nl : 0) + nl + nd * npos + 2 * npos + 2 * c_.K + 2 * nl + 4) * c_.N * 8;
    }
    size_t footprint(const Job& j) const {
        const int l = j.a->level;
        if (j.kind == JOB_LOGREG) return 6 * ct_bytes() + ks_bytes(l, true);
        return ks_bytes(l, j.kind == JOB_MULT);
    }

    void admit() {
        while (!queue_.empty() && (int)active_.size() < max_slots_ && used_ + queue_.front()->bytes <= budget_) {
            std::unique_ptr<Job> j = std::move(queue_.front());
            queue_.pop_front();
            if (j->kind == JOB_LOGREG)
                for (int i = 0; i < 6; ++i) j->tmp.push_back(ct_create(c_));
            used_ += j->bytes;
            active_.push_back(std::move(j));
        }
    }

    const KeySwitchingKey& rkey(int index, uint32_t& kinv) {
        const uint32_t k = c_.rot_index(index);
        auto it = c_.rot_keys.find(k);
        if (it == c_.rot_keys.end()) throw Error(E_NOKEY, "no rotation key for index " + std::to_string(index));
        kinv = c_.auto_inverse(k);
        return *it->second;
    }
    KsItem mult_item(const Ciphertext& a, const Ciphertext& b, Ciphertext& out) {
        if (!c_.eval_key) throw Error(E_NOKEY, "no relinearization key loaded");
        return KsItem{true, &a, &b, &out, c_.eval_key.get(), 0, true};
    }
    KsItem rot_item(const Ciphertext& a, int index, Ciphertext& out) {
        uint32_t kinv;
        const KeySwitchingKey& k = rkey(index, kinv);
        return KsItem{false, &a, nullptr, &out, &k, kinv, false};
    }

    // Local ops before this step's key switches, and the key-switched ops.
    void pre(Job& j, std::vector<KsItem>& items) {
        if (j.kind == JOB_MULT) {
            items.push_back(mult_item(*j.a, *j.b, *j.out));
            return;
        }
        if (j.kind == JOB_ROTATE) {
            if (c_.rot_index(j.index) == 1) {
                ct_copy(c_, *j.a, *j.out);
                return;
            }
            items.push_back(rot_item(*j.a, j.index, *j.out));
            return;
        }
        Ciphertext &y = *j.tmp[0], &t = *j.tmp[1], &y2 = *j.tmp[2], &u = *j.tmp[3], &y3 = *j.tmp[4],
                   &v = *j.tmp[5];
        if (j.step == 0) eval_mult_plain(c_, *j.a, *j.w, y, true);
        if (j.step < j.n_rot) {
            items.push_back(rot_item(y, 1 << j.step, t));
        } else if (j.step == j.n_rot) {
            eval_mult_scalar(c_, y, j.c3, u, true);
            eval_mult_scalar(c_, y, j.c1, v, true);
            items.push_back(mult_item(y, y, y2));
        } else {
            items.push_back(mult_item(y2, u, y3));
        }
    }

    // Local ops after this step's key switches; true when the job is done.
    bool post(Job& j) {
        if (j.kind != JOB_LOGREG) return true;
        Ciphertext &y = *j.tmp[0], &t = *j.tmp[1], &y3 = *j.tmp[4], &v = *j.tmp[5];
        if (j.step < j.n_rot) {
            eval_add(c_, y, t, y);
        } else if (j.step == j.n_rot + 1) {
            drop_to_level(c_, v, y3.level);
            eval_add(c_, y3, v, *j.out);
            eval_add_scalar(c_, *j.out, j.c0, *j.out);
            return true;
        }
        ++j.step;
        return false;
    }

    Context& c_;
    int max_slots_;
    uint64_t budget_, used_ = 0;
    int64_t next_id_ = 1;
    // jobs finished in one step, released when the event recorded after them fires
    struct Completion {
        hipEvent_t ev = nullptr;
        std::vector<std::unique_ptr<Job>> jobs;
    };
    std::deque<std::unique_ptr<Job>> queue_;
    std::vector<std::unique_ptr<Job>> active_;
    std::deque<Completion> events_;
};

Server* server_create(Context& c, int max_slots, uint64_t budget) {
    if (max_slots < 1) throw Error(E_ARG, "max_slots must be positive");
    return new Server(c, max_slots, budget);
}
void server_destroy(Server* s) { delete s; }

int64_t server_submit_mult(Server& s, const Ciphertext& a, const Ciphertext& b, Ciphertext& out) {
    if (a.level < 0 || b.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (&out == &a || &out == &b) throw Error(E_ARG, "server job output must not alias its inputs");
    auto j = std::make_unique<Job>();
    j->kind = JOB_MULT;
    j->a = &a;
    j->b = &b;
    j->out = &out;
    return s.submit(std::move(j));
}
int64_t server_submit_rotate(Server& s, const Ciphertext& a, int index, Ciphertext& out) {
    if (a.level < 0) throw Error(E_STATE, "ciphertext not loaded");
    if (&out == &a) throw Error(E_ARG, "server job output must not alias its inputs");
    auto j = std::make_unique<Job>();
    j->kind = JOB_ROTATE;
    j->a = &a;
    j->index = index;
    j->out = &out;
    return s.submit(std::move(j));
}
int64_t server_submit_logreg(Server& s, const Ciphertext& x, const Plaintext& w, int n_rot, double c0, double c1,
                             double c3, Ciphertext& out) {
    if (x.level < 0 || w.level < 0) throw Error(E_STATE, "input not loaded");
    if (&out == &x) throw Error(E_ARG, "server job output must not alias its inputs");
    if (n_rot < 0 || n_rot > 30) throw Error(E_ARG, "n_rot out of range");
    if (x.level < 3) throw Error(E_LEVEL, "logistic scoring needs level 3 or above (three rescales)");
    if (x.ctx->scaling_technique != FIXEDMANUAL) throw Error(E_STATE, "logistic scoring jobs run under FIXEDMANUAL");
    auto j = std::make_unique<Job>();
    j->kind = JOB_LOGREG;
    j->a = &x;
    j->w = &w;
    j->n_rot = n_rot;
    j->c0 = c0;
    j->c1 = c1;
    j->c3 = c3;
    j->out = &out;
    return s.submit(std::move(j));
}
int server_step(Server& s) { return s.step(); }
std::vector<int64_t> server_poll(Server& s) { return s.poll(); }
void server_counts(const Server& s, int& active, int& queued, uint64_t& used) {
    active = s.active();
    queued = s.queued();
    used = s.used_bytes();
}

}  // namespace fides
